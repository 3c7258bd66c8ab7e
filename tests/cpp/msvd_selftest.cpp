// Self-test of the C++ adapter (include/ofrr/msvd.hpp): the reference's
// container and solver interfaces driven through libofrr.so on the GPU.
// Checks mirror the reference's own tests: the mvops_test.cpp KATs (gemm
// 3x2*2x4, iamax planted maxima, syev/sygv tridiagonal spectra), the
// hessenberg_test.cpp tolerance and skip pattern, and closed-form Laplacian
// spectra for the eigensolver drivers.  Prints one line per check; exit code 0
// when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "ofrr/msvd.hpp"

using namespace msvd;

static int g_fail = 0;
static void check(bool ok, const char *name, double detail = 0.0) {
   std::printf("%s %s (%.3e)\n", ok ? "PASS" : "FAIL", name, detail);
   std::fflush(stdout);
   if (!ok) ++g_fail;
}

template <typename T> static std::vector<T> host_of(const Matrix<T> &A) {
   std::vector<T> h(A.ld() * A.cols());
   (void)hipMemcpy(h.data(), A.data(), h.size() * sizeof(T), hipMemcpyDefault);
   return h;
}

int main() {
   // ---- matrix input: the bcsstk01 fixture in the binary CSR format
   {
      const char *root = std::getenv("OFRR_FIXTURES");
      const std::string pre = std::string(root ? root : "tests/golden/suitesparse") + "/bcsstk01";
      CSRMatrix m = MatrixReader::read_csr_matrix(pre);
      check(m.rows == 48 && m.cols == 48 && m.row_ptr.size() == 49 && m.values.size() == 400 &&
                m.row_ptr.back() == 400,
            "MatrixReader::read_csr_matrix bcsstk01");
      bool thrown = false;
      try {
         MatrixReader::read_csr_matrix(pre + "_missing");
      } catch (const std::runtime_error &e) {
         thrown = std::string(e.what()).find("Failed to read row pointers from") == 0;
      }
      check(thrown, "MatrixReader missing file message");
   }
   // ---- containers: copy is deep, move empties the source, round trips
   {
      Vector<double> v(1000);
      v.fill_random(3);
      Vector<double> w = v;
      w.to_host();
      Vector<double> u(std::move(v));
      check(v.data() == nullptr && v.length() == 0 && u.length() == 1000, "vector move");
      u.to_host();
      bool same = true;
      for (size_t i = 0; i < 1000; ++i) same = same && (u[i] == w[i]);
      check(same && u.location() == Location::kHOST, "vector copy / transfer");
   }
   // ---- gemm KAT (mvops_test.cpp:455-721): C = A B, A 3x2, B 2x4
   {
      Matrix<double> A(3, 2, Location::kHOST), B(2, 4, Location::kHOST), C(3, 4, Location::kHOST);
      const double a[3][2] = {{1, 3}, {2, 4}, {5, 6}};
      const double b[2][4] = {{1, 2, 3, 4}, {5, 6, 7, 8}};
      const double c[3][4] = {{16, 20, 24, 28}, {22, 28, 34, 40}, {35, 46, 57, 68}};
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 2; ++j) A(i, j) = a[i][j];
      for (int i = 0; i < 2; ++i) for (int j = 0; j < 4; ++j) B(i, j) = b[i][j];
      MSVDStatus st = gemm<double, double, double>(false, false, 1.0, A, B, 0.0, C);
      auto h = host_of(C);
      double err = 0;
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) err = std::fmax(err, std::fabs(h[i + 3 * j] - c[i][j]));
      check(st == MSVDStatus::kSuccess && err == 0.0, "gemm 3x2*2x4 KAT", err);
      // A^T A (transA): 2x2 = [[30, 41], [41, 61]]
      Matrix<double> G(2, 2, Location::kHOST);
      st = gemm<double, double, double>(true, false, 1.0, A, A, 0.0, G);
      check(st == MSVDStatus::kSuccess && G(0, 0) == 30 && G(0, 1) == 41 && G(1, 1) == 61, "gemm transA");
   }
   // ---- iamax KAT (mvops_test.cpp:774-1150): x_i = (i - 65536) / 131072, planted -8 at 72345
   {
      const int n = 131072;
      std::vector<double> h(n);
      for (int i = 0; i < n; ++i) h[i] = (i - 65536) / 131072.0;
      h[72345] = -8.0;
      Vector<double> x(n);
      (void)hipMemcpy(x.data(), h.data(), n * 8, hipMemcpyHostToDevice);
      int idx = -1;
      double val = 0;
      MSVDStatus st = iamax<double, double>(x, idx, val);
      check(st == MSVDStatus::kSuccess && idx == 72346 && val == -8.0, "iamax planted KAT (1-based)", idx);
   }
   // ---- syev / sygv KATs: tridiag(-1, 2, -1), n = 10; B = diag(1..10)
   {
      const int n = 10;
      Matrix<double> A(n, n, Location::kHOST), B(n, n, Location::kHOST);
      for (int i = 0; i < n; ++i) {
         A(i, i) = 2.0;
         B(i, i) = i + 1.0;
         if (i + 1 < n) A(i, i + 1) = A(i + 1, i) = -1.0;
      }
      Matrix<double> A2 = A;
      Vector<double> W(n, Location::kHOST), W2(n, Location::kHOST);
      double lw = 0;
      check(syev(A, W, &lw, -1) == MSVDStatus::kSuccess, "syev workspace query");
      MSVDStatus st = syev<double>(A, W, nullptr, 0);
      double err = 0;
      for (int i = 0; i < n; ++i) err = std::fmax(err, std::fabs(W[i] - (2.0 - 2.0 * std::cos((i + 1) * M_PI / (n + 1)))));
      check(st == MSVDStatus::kSuccess && err < 1e-13, "syev tridiag(-1,2,-1) KAT", err);
      st = sygv<double>(A2, B, W2, nullptr, 0);
      check(st == MSVDStatus::kSuccess && std::fabs(W2[0] - 0.0141309446511971) < 1e-12 &&
                std::fabs(W2[n - 1] - 2.38801544566837) < 1e-12,
            "sygv diag(1..10) KAT", W2[0]);
   }
   // ---- hessenberg (hessenberg_test.cpp): 2000 x 200 uniform, ||A - QR||_max < 1e-10;
   // duplicated columns 3..5 are dropped
   {
      const int m = 2000, n = 200;
      Matrix<double> A(m, n), Q(m, n), R(n, n, Location::kHOST);
      A.fill_random(7);
      {
         auto h = host_of(A);
         for (int j = 3; j <= 5; ++j)
            for (int i = 0; i < m; ++i) h[i + (size_t)j * m] = h[i + 2 * (size_t)m];
         (void)hipMemcpy(A.data(), h.data(), h.size() * 8, hipMemcpyHostToDevice);
      }
      std::vector<int> skip = hessenberg<double, double, double>(A, Q, R, 1e-12);
      auto ha = host_of(A), hq = host_of(Q);
      double err = 0;
      for (int j = 0; j < n; ++j)
         for (int i = 0; i < m; i += 7) {
            double s = 0;
            for (int p = 0; p <= j; ++p) if (!skip[p]) s += hq[i + (size_t)p * m] * R(p, j);
            if (!skip[j]) err = std::fmax(err, std::fabs(s - ha[i + (size_t)j * m]));
         }
      bool pattern = skip[3] == 1 && skip[4] == 1 && skip[5] == 1 && skip[2] == 0 && skip[6] == 0;
      check(err < 1e-10 && pattern, "hessenberg A = QR and skip pattern", err);
      bool threw = false;
      try {
         Matrix<double> Ah(m, n, Location::kHOST);
         hessenberg<double, double, double>(Ah, Q, R);
      } catch (const std::runtime_error &) {
         threw = true;
      }
      check(threw, "hessenberg rejects a host A");
   }
   // ---- Gram-Schmidt QR (qr_test.cpp:30-651): 2000 x 200 uniform, Q^T Q = I and
   // A = QR within 1e-10 (FP64) / 2e-6 (FP32 cgs2); duplicated columns 3..5 dropped;
   // the column steps reproduce the factorization
   {
      const int m = 2000, n = 200;
      Matrix<double> A(m, n), Q(m, n), R(n, n, Location::kHOST);
      A.fill_random(11);
      auto qr_err = [&](const std::vector<double> &ha, const std::vector<double> &hq,
                        const Matrix<double> &Rm, const std::vector<int> &skip) {
         double e = 0;
         for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; i += 7) {
               double s = 0;
               for (int p = 0; p <= j; ++p) s += hq[i + (size_t)p * m] * Rm(p, j);
               e = std::fmax(e, std::fabs(s - ha[i + (size_t)j * m]));
            }
         for (int a = 0; a < n; a += 3)
            for (int b = a; b < n; b += 5) {
               if (skip[a] || skip[b]) continue;
               double s = 0;
               for (int i = 0; i < m; ++i) s += hq[i + (size_t)a * m] * hq[i + (size_t)b * m];
               e = std::fmax(e, std::fabs(s - (a == b ? 1.0 : 0.0)));
            }
         return e;
      };
      std::vector<int> skip = mgs<double, double, double>(A, Q, R);
      double err = qr_err(host_of(A), host_of(Q), R, skip);
      int dropped = 0;
      for (int s : skip) dropped += s;
      check(err < 1e-10 && dropped == 0, "mgs Q^T Q = I and A = QR", err);
      skip = cgs2<double, double, double>(A, Q, R);
      err = qr_err(host_of(A), host_of(Q), R, skip);
      check(err < 1e-10, "cgs2 Q^T Q = I and A = QR", err);
      {
         auto h = host_of(A);
         for (int j = 3; j <= 5; ++j)
            for (int i = 0; i < m; ++i) h[i + (size_t)j * m] = h[i + (size_t)(j - 3) * m];
         (void)hipMemcpy(A.data(), h.data(), h.size() * 8, hipMemcpyHostToDevice);
      }
      for (int v = 0; v < 3; ++v) {
         const char *name[3] = {"mgs", "mgs_v2", "cgs"};
         skip = v == 0 ? mgs<double, double, double>(A, Q, R, 1e-12, 1e-12, 0.70710678118654752)
                : v == 1 ? mgs_v2<double, double, double>(A, Q, R, 1e-12)
                         : cgs<double, double, double>(A, Q, R, 1e-12);
         err = qr_err(host_of(A), host_of(Q), R, skip);
         bool pattern = skip[3] == 1 && skip[4] == 1 && skip[5] == 1 && skip[2] == 0 && skip[6] == 0;
         check(err < 1e-10 && pattern, (std::string(name[v]) + " duplicated columns dropped").c_str(), err);
      }
      // column steps on the same matrix: mgs_step into a host R, cgs_step into R_device + r_diag
      Matrix<double> Qs(m, n), Rs(n, n, Location::kHOST), Rd(n, n);
      Vector<double> rdiag(n, Location::kHOST);
      std::vector<int> ss(n, 0), sref = mgs<double, double, double>(A, Q, R, 1e-12);
      (void)hipMemcpy(Qs.data(), A.data(), (size_t)m * n * 8, hipMemcpyDeviceToDevice);
      for (int j = 0; j < n; ++j)
         for (int i = 0; i < n; ++i) Rs(i, j) = 0.0;
      for (int i = 0; i < n; ++i) mgs_step<double, double>(Qs, Rs, ss, i, 1e-12);
      double dq = 0;
      auto hq = host_of(Q), hqs = host_of(Qs);
      for (size_t t = 0; t < hq.size(); ++t) dq = std::fmax(dq, std::fabs(hq[t] - hqs[t]));
      for (int j = 0; j < n; ++j)
         for (int i = 0; i < n; ++i) dq = std::fmax(dq, std::fabs(R(i, j) - Rs(i, j)));
      check(dq < 1e-12 && ss == sref, "mgs_step sequence = mgs", dq);
      sref = cgs<double, double, double>(A, Q, R, 1e-12);
      (void)hipMemcpy(Qs.data(), A.data(), (size_t)m * n * 8, hipMemcpyDeviceToDevice);
      std::fill(ss.begin(), ss.end(), 0);
      for (int i = 0; i < n; ++i) rdiag[i] = 1.0;
      for (int i = 0; i < n; ++i) cgs_step<double, double>(Qs, Rd, rdiag, ss, i, 1e-12);
      auto hrd = host_of(Rd);
      dq = 0;
      hq = host_of(Q);
      hqs = host_of(Qs);
      for (size_t t = 0; t < hq.size(); ++t) dq = std::fmax(dq, std::fabs(hq[t] - hqs[t]));
      for (int j = 0; j < n; ++j) {
         for (int i = 0; i < j; ++i) dq = std::fmax(dq, std::fabs(R(i, j) - hrd[i + (size_t)j * n]));
         dq = std::fmax(dq, std::fabs(R(j, j) - rdiag[j]));
      }
      check(dq == 0.0 && ss == sref, "cgs_step sequence = cgs (bitwise)", dq);
      // FP32 cgs2 (qr_test.cpp:386-470): 2e-6
      Matrix<float> Af(m, n), Qf(m, n), Rf(n, n, Location::kHOST);
      Af.fill_random(12);
      std::vector<int> sf = cgs2<float, float, float>(Af, Qf, Rf);
      auto haf = host_of(Af), hqf = host_of(Qf);
      double ef = 0;
      for (int a = 0; a < n; a += 3)
         for (int b = a; b < n; b += 5) {
            double s = 0;
            for (int i = 0; i < m; ++i) s += (double)hqf[i + (size_t)a * m] * hqf[i + (size_t)b * m];
            ef = std::fmax(ef, std::fabs(s - (a == b ? 1.0 : 0.0)));
         }
      for (int j = 0; j < n; j += 3)
         for (int i = 0; i < m; i += 11) {
            double s = 0;
            for (int p = 0; p <= j; ++p) s += (double)hqf[i + (size_t)p * m] * Rf(p, j);
            ef = std::fmax(ef, std::fabs(s - haf[i + (size_t)j * m]));
         }
      check(ef < 2e-6 && sf[0] == 0, "cgs2 FP32 within 2e-6", ef);
      bool threw = false;
      try {
         Matrix<double> Rdev(n, n);
         mgs<double, double, double>(A, Q, Rdev);
      } catch (const std::runtime_error &e) {
         threw = std::string(e.what()) == "Output matrix R must be on host";
      }
      check(threw, "mgs rejects a device R with the reference's message");
   }
   // ---- si_eig / arnoldi_eig on the dense 2D Laplacian, 24 x 24 grid
   {
      const int N = 24, n = N * N;
      Matrix<double> A(n, n, Location::kHOST);
      for (int y = 0; y < N; ++y)
         for (int x = 0; x < N; ++x) {
            const int i = x + N * y;
            A(i, i) = 4.0;
            if (x > 0) A(i, i - 1) = -1.0;
            if (x + 1 < N) A(i, i + 1) = -1.0;
            if (y > 0) A(i, i - N) = -1.0;
            if (y + 1 < N) A(i, i + N) = -1.0;
         }
      std::vector<double> ex;
      for (int p = 1; p <= N; ++p)
         for (int q = 1; q <= N; ++q)
            ex.push_back(4 * std::pow(std::sin(p * M_PI / (2.0 * (N + 1))), 2) + 4 * std::pow(std::sin(q * M_PI / (2.0 * (N + 1))), 2));
      std::sort(ex.begin(), ex.end(), [](double a, double b) { return a > b; });
      Matrix<double> V(n, 32), D(32, 32, Location::kHOST);
      V.fill_random(1);
      MSVDStatus st = si_eig<double, double, double>(A, V, D, 200, 1e-10, false, 1e-12);
      double err = 0;
      for (int i = 0; i < 16; ++i) err = std::fmax(err, std::fabs(D(i, i) - ex[i]) / ex[i]);
      check(st == MSVDStatus::kSuccess && err < 1e-12, "si_eig dense 2D Laplacian top 16", err);
      Matrix<double> Va(n, 45), Da(15, 15, Location::kHOST);
      // restarted Arnoldi on a clustered top spectrum: the restart count, not
      // the arithmetic, bounds the accuracy (2e-7 after 40 restarts)
      st = arnoldi_eig<double, double, double>(A, Va, Da, 100, 1e-10, true, 1e-12);
      err = std::fabs(Da(0, 0) - ex[0]) / ex[0];
      check(st == MSVDStatus::kSuccess && err < 1e-6, "arnoldi_eig dense 2D Laplacian top", err);
      // rr_eig on exact eigenvectors' span: the Ritz values are eigenvalues
      Matrix<double> Vr(n, 16);
      (void)hipMemcpy2D(Vr.data(), n * 8, V.data(), n * 8, n * 8, 16, hipMemcpyDeviceToDevice);
      Matrix<double> Dr(16, 16, Location::kHOST);
      st = rr_eig<double, double, double>(A, Vr, Dr, false, 1e-12);
      err = 0;
      for (int i = 0; i < 8; ++i) err = std::fmax(err, std::fabs(Dr(i, i) - ex[i]) / ex[i]);
      check(st == MSVDStatus::kSuccess && err < 1e-10, "rr_eig on a converged basis", err);
   }
   // ---- si_svd / rr_svd: A = U diag(s) W^T, s_i = 2^-i/4, 4096 x 256 (FP32)
   {
      const int M = 4096, N = 256, k = 32;
      Matrix<float> A(M, N);
      {
         Matrix<double> Ur(M, N), Wr(N, N);
         Ur.fill_random(11);
         Wr.fill_random(12);
         // orthonormalise on the host (modified Gram-Schmidt) and form A
         auto u = host_of(Ur), w = host_of(Wr);
         auto mgs = [](std::vector<double> &Q, int rows, int cols) {
            for (int j = 0; j < cols; ++j) {
               for (int p = 0; p < j; ++p) {
                  double d = 0;
                  for (int i = 0; i < rows; ++i) d += Q[i + (size_t)p * rows] * Q[i + (size_t)j * rows];
                  for (int i = 0; i < rows; ++i) Q[i + (size_t)j * rows] -= d * Q[i + (size_t)p * rows];
               }
               double nr = 0;
               for (int i = 0; i < rows; ++i) nr += Q[i + (size_t)j * rows] * Q[i + (size_t)j * rows];
               nr = std::sqrt(nr);
               for (int i = 0; i < rows; ++i) Q[i + (size_t)j * rows] /= nr;
            }
         };
         mgs(u, M, 64);
         mgs(w, N, 64);
         std::vector<float> a((size_t)M * N, 0.f);
         for (int p = 0; p < 64; ++p) {
            const double s = std::pow(2.0, -p / 4.0);
            for (int j = 0; j < N; ++j)
               for (int i = 0; i < M; ++i) a[i + (size_t)j * M] += (float)(u[i + (size_t)p * M] * s * w[j + (size_t)p * N]);
         }
         (void)hipMemcpy(A.data(), a.data(), a.size() * 4, hipMemcpyHostToDevice);
      }
      Matrix<double> U(M, k), S(k, k, Location::kHOST), V(N, k);
      V.fill_random(5);
      MSVDStatus st = si_svd<float, double, double>(A, U, S, V, 100, 1e-8, false, 1e-12);
      double err = 0;
      for (int i = 0; i < 16; ++i) err = std::fmax(err, std::fabs(S(i, i) - std::pow(2.0, -i / 4.0)) / std::pow(2.0, -i / 4.0));
      check(st == MSVDStatus::kSuccess && err < 1e-5, "si_svd low-rank top 16 (FP32 matrix)", err);
      Matrix<double> S2(k, k, Location::kHOST);
      st = rr_svd<float, double, double>(A, U, S2, V, false, 1e-12);
      double err2 = 0;
      for (int i = 0; i < 16; ++i) err2 = std::fmax(err2, std::fabs(S2(i, i) - S(i, i)) / S(i, i));
      check(st == MSVDStatus::kSuccess && err2 < 1e-10, "rr_svd on the converged basis", err2);
   }
   // ---- gemm: transposes and the reference's exceptions (mvops_test.cpp:622-769)
   {
      Matrix<double> A(2, 3, Location::kHOST), B(4, 2, Location::kHOST), C(3, 4, Location::kHOST);
      const double at[2][3] = {{1, 2, 5}, {3, 4, 6}};
      const double bt[4][2] = {{1, 5}, {2, 6}, {3, 7}, {4, 8}};
      const double c[3][4] = {{16, 20, 24, 28}, {22, 28, 34, 40}, {35, 46, 57, 68}};
      for (int i = 0; i < 2; ++i) for (int j = 0; j < 3; ++j) A(i, j) = at[i][j];
      for (int i = 0; i < 4; ++i) for (int j = 0; j < 2; ++j) B(i, j) = bt[i][j];
      C.fill(0.0);
      MSVDStatus st = gemm<double, double, double>(true, true, 1.0, A, B, 0.0, C);
      bool same = st == MSVDStatus::kSuccess;
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) same = same && C(i, j) == c[i][j];
      check(same, "gemm A^T B^T KAT (host matrices, exact)");
      // beta on the device: C = 2 A^T B^T - C
      Matrix<double> Ad = A, Bd = B, Cd = C;
      Ad.to_device(); Bd.to_device(); Cd.to_device();
      st = gemm<double, double, double>(true, true, 2.0, Ad, Bd, -1.0, Cd);
      Cd.to_host();
      same = st == MSVDStatus::kSuccess;
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) same = same && Cd(i, j) == c[i][j];
      check(same, "gemm beta on the device (no host round trip)");
      auto msg = [](auto &&f) {
         try { f(); } catch (const std::runtime_error &e) { return std::string(e.what()); }
         return std::string("no exception");
      };
      Matrix<float> Af(3, 2, Location::kHOST), Bf(2, 4, Location::kDEVICE), Cf(3, 4, Location::kHOST);
      check(msg([&] { gemm<float, float, float>(false, false, 1.0f, Af, Bf, 0.0f, Cf); }) ==
                "All matrices must be in the same location (CPU or GPU)", "gemm location message");
      Matrix<float> A2(3, 3, Location::kHOST), B2(2, 4, Location::kHOST), C2(3, 4, Location::kHOST);
      check(msg([&] { gemm<float, float, float>(false, false, 1.0f, A2, B2, 0.0f, C2); }) ==
                "Matrix dimensions do not match for multiplication", "gemm dimension message");
      Matrix<float> A3(3, 2, Location::kHOST), B3(2, 4, Location::kHOST), C3(3, 3, Location::kHOST);
      check(msg([&] { gemm<float, float, float>(false, false, 1.0f, A3, B3, 0.0f, C3); }) ==
                "Output matrix C has wrong dimensions", "gemm output message");
      Matrix<double> Ah(3, 2, Location::kHOST), Bh(2, 4, Location::kHOST);
      Matrix<float> Ch(3, 4, Location::kHOST);
      check(msg([&] { gemm<double, float, double>(false, false, 1.0, Ah, Bh, 0.0, Ch); }) ==
                "For CPU GEMM, input, output, and compute types must be the same", "gemm CPU mixed-type message");
      // a general-size product against a host FP64 loop (both transposes)
      const int m = 300, n = 170, kk = 211;
      Matrix<double> P(kk, m), Qm(n, kk), G(m, n);
      P.fill_random(21); Qm.fill_random(22);
      st = gemm<double, double, double>(true, true, 1.0, P, Qm, 0.0, G);
      auto hp = host_of(P), hq = host_of(Qm), hg = host_of(G);
      double err = 0;
      for (int i = 0; i < m; i += 7)
         for (int j = 0; j < n; j += 5) {
            double s2 = 0;
            for (int p = 0; p < kk; ++p) s2 += hp[p + (size_t)i * kk] * hq[j + (size_t)p * n];
            err = std::fmax(err, std::fabs(s2 - hg[i + (size_t)j * m]));
         }
      check(st == MSVDStatus::kSuccess && err < 1e-12, "gemm 300x170x211 A^T B^T", err);
   }
   // ---- containers: scale, cast, submatrix, fill_random range
   {
      Matrix<double> A(50, 40);
      A.fill_random(3);
      auto h0 = host_of(A);
      double lo = 1, hi = -1;
      for (double v : h0) { lo = std::fmin(lo, v); hi = std::fmax(hi, v); }
      check(lo >= -1.0 && hi < 1.0 && lo < -0.9 && hi > 0.9, "fill_random uniform in [-1, 1)", lo);
      Matrix<double> S = A.submatrix(10, 5, 20, 30);
      S.scale(-2.0);
      auto h1 = host_of(A);
      bool ok = S.ld() == 50 && S.rows() == 20 && S.cols() == 30 && S.location() == Location::kDEVICE;
      for (int j = 0; j < 40; ++j)
         for (int i = 0; i < 50; ++i) {
            const bool in = i >= 10 && i < 30 && j >= 5 && j < 35;
            ok = ok && h1[i + 50 * j] == (in ? -2.0 * h0[i + 50 * j] : h0[i + 50 * j]);
         }
      check(ok, "submatrix view + scale on the device");
      Matrix<float> F = A.cast<float>();
      auto hf = host_of(F);
      ok = F.rows() == 50 && F.cols() == 40 && F.location() == Location::kDEVICE;
      for (size_t t = 0; t < hf.size(); ++t) ok = ok && hf[t] == (float)h1[t];
      check(ok, "Matrix::cast<float> on the device");
      Vector<double> v(100, Location::kHOST);
      for (int i = 0; i < 100; ++i) v[i] = i;
      v.scale(0.5);
      Vector<float> vf = v.cast<float>();
      check(v[7] == 3.5 && vf[9] == 4.5f && vf.location() == Location::kHOST, "Vector::scale / cast on the host");
      bool threw = false;
      try { A.submatrix(45, 0, 10, 1); } catch (const std::runtime_error &) { threw = true; }
      check(threw, "submatrix out of bounds throws");
   }
   // ---- hessenberg_v2 / v3 / v3_step (hessenberg_test.cpp: 2000 x 200, 1e-10)
   {
      const int m = 2000, n = 200;
      Matrix<double> A(m, n), Q2(m, n), Q3(m, n), Qs(m, n), R2(n, n, Location::kHOST), R3(n, n, Location::kHOST);
      A.fill_random(17);
      {
         auto h = host_of(A);
         for (int j = 3; j <= 5; ++j)
            for (int i = 0; i < m; ++i) h[i + (size_t)j * m] = h[i + (size_t)(j - 3) * m];
         (void)hipMemcpy(A.data(), h.data(), h.size() * 8, hipMemcpyHostToDevice);
      }
      std::vector<int> s2 = hessenberg_v2<double, double, double>(A, Q2, R2, 1e-12);
      std::vector<int> s3 = hessenberg_v3<double, double, double>(A, Q3, R3, 1e-12);
      bool pattern = s2[3] && s2[4] && s2[5] && !s2[2] && !s2[6] && s3 == s2;
      // v3 reproduced column by column with the step entry point
      (void)hipMemcpy(Qs.data(), A.data(), (size_t)m * n * 8, hipMemcpyDeviceToDevice);
      Matrix<double> Rdummy(n, n, Location::kHOST), sf(n, n);
      Vector<int> perm(n), skipd(n);
      int skiph = 0;
      std::vector<int> ss(n);
      for (int i = 0; i < n; ++i)
         ss[i] = hessenberg_v3_step<double, double>(Qs, Rdummy, &skiph, skipd.data(), i, perm.data(), sf,
                                                    (double *)nullptr, 1e-12);
      auto hq3 = host_of(Q3), hqs = host_of(Qs), hsf = host_of(sf);
      bool same = ss == s3;
      for (size_t t = 0; t < hq3.size(); ++t) same = same && hq3[t] == hqs[t];
      for (int j = 0; j < n; ++j)
         for (int i = 0; i <= j; ++i) same = same && R3(i, j) == hsf[i + (size_t)j * n];
      auto ha = host_of(A), hq2 = host_of(Q2);
      double err = 0;
      for (int j = 0; j < n; ++j)
         for (int i = 0; i < m; i += 7) {
            double s = 0;
            for (int p = 0; p <= j; ++p) if (!s2[p]) s += hq2[i + (size_t)p * m] * R2(p, j);
            if (!s2[j]) err = std::fmax(err, std::fabs(s - ha[i + (size_t)j * m]));
         }
      check(err < 1e-10 && pattern, "hessenberg_v2 A = QR and skip pattern, v3 same drops", err);
      check(same, "hessenberg_v3_step sequence = hessenberg_v3 (bitwise)");
   }
#ifdef OFRR_MSVD_HALF
   // ---- half precision (hessenberg_test.cpp:161-224, mvops_test.cpp:455-721)
   {
      Matrix<__half> A(3, 2, Location::kHOST), B(2, 4, Location::kHOST), C(3, 4, Location::kDEVICE);
      const float a[3][2] = {{1, 3}, {2, 4}, {5, 6}};
      const float b[2][4] = {{1, 2, 3, 4}, {5, 6, 7, 8}};
      const float c[3][4] = {{16, 20, 24, 28}, {22, 28, 34, 40}, {35, 46, 57, 68}};
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 2; ++j) A(i, j) = __float2half(a[i][j]);
      for (int i = 0; i < 2; ++i) for (int j = 0; j < 4; ++j) B(i, j) = __float2half(b[i][j]);
      A.to_device(); B.to_device();
      C.fill(__float2half(0.0f));
      MSVDStatus st = gemm<__half, __half, float>(false, false, 1.0f, A, B, 0.0f, C);
      C.to_host();
      float e1 = 0;
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) e1 = std::fmax(e1, std::fabs(__half2float(C(i, j)) - c[i][j]));
      check(st == MSVDStatus::kSuccess && e1 <= 0.1f, "gemm<__half,__half,float> KAT", e1);
      Matrix<__half> C2(3, 4, Location::kDEVICE);
      st = gemm<__half, __half, __half>(false, false, __float2half(1.0f), A, B, __float2half(0.0f), C2);
      C2.to_host();
      float e2 = 0;
      for (int i = 0; i < 3; ++i) for (int j = 0; j < 4; ++j) e2 = std::fmax(e2, std::fabs(__half2float(C2(i, j)) - c[i][j]));
      check(st == MSVDStatus::kSuccess && e2 <= 0.2f, "gemm<__half,__half,__half> KAT", e2);
   }
   {
      const size_t m = 2000, n = 100;
      Matrix<__half> A(m, n), Q(m, n), R(n, n, Location::kHOST);
      A.fill_random();
      std::vector<int> skip = hessenberg<__half, __half, float>(A, Q, R);
      bool clean = check_special_values(Q) == MSVDStatus::kSuccess;
      Matrix<__half> QR(m, n);
      R.to_device();
      clean = clean && check_special_values(R) == MSVDStatus::kSuccess;
      gemm<__half, __half, float>(false, false, 1.0f, Q, R, 0.0f, QR);
      A.to_host();
      QR.to_host();
      float md = 0;
      for (size_t j = 0; j < n; ++j)
         for (size_t i = 0; i < m; ++i) md = std::fmax(md, std::fabs(__half2float(A(i, j)) - __half2float(QR(i, j))));
      check(clean && md < 5e-2f, "hessenberg<__half,__half,float> 2000x100: A = QR within 5e-2", md);
      A.fill_random();
      A.to_device();
      R.to_host();
      (void)hipMemcpy(&A(0, 3), &A(0, 0), 3 * sizeof(__half) * m, hipMemcpyDeviceToDevice);
      std::vector<int> skip2 = hessenberg<__half, __half, float>(A, Q, R, __float2half(1e-2f));
      bool pattern = true;
      for (size_t i = 0; i < n; ++i) pattern = pattern && skip2[i] == ((i == 3 || i == 4 || i == 5) ? 1 : 0);
      check(pattern, "hessenberg<__half> duplicated columns 3..5 dropped at 1e-2");
      std::vector<int> skip3 = hessenberg_v3<__half, __half, float>(A, Q, R, __float2half(1e-2f));
      check(skip3 == skip2, "hessenberg_v3<__half> same drops");
   }
   {
      // ofrr_eig / rr_eig / si_eig with half-precision operands on the 2D Laplacian 16 x 16
      const int N = 16, n = N * N;
      Matrix<double> A(n, n, Location::kHOST);
      for (int y = 0; y < N; ++y)
         for (int x = 0; x < N; ++x) {
            const int i = x + N * y;
            A(i, i) = 4.0;
            if (x > 0) A(i, i - 1) = -1.0;
            if (x + 1 < N) A(i, i + 1) = -1.0;
            if (y > 0) A(i, i - N) = -1.0;
            if (y + 1 < N) A(i, i + N) = -1.0;
         }
      std::vector<double> ex;
      for (int p = 1; p <= N; ++p)
         for (int q = 1; q <= N; ++q)
            ex.push_back(4 * std::pow(std::sin(p * M_PI / (2.0 * (N + 1))), 2) + 4 * std::pow(std::sin(q * M_PI / (2.0 * (N + 1))), 2));
      std::sort(ex.begin(), ex.end(), [](double a, double b) { return a > b; });
      Matrix<__half> Ah = A.cast<__half>();  // the stencil is exact in FP16
      Ah.to_device();
      Matrix<__half> V(n, 24), D(24, 24, Location::kHOST);
      V.fill_random(9);
      MSVDStatus st = si_eig<__half, __half, float>(Ah, V, D, 200, 1e-8f, false, 1e-12f);
      double err = 0;
      for (int i = 0; i < 12; ++i) err = std::fmax(err, std::fabs((double)__half2float(D(i, i)) - ex[i]) / ex[i]);
      check(st == MSVDStatus::kSuccess && err < 1e-3, "si_eig<__half,__half,float> (FP16 growth blocks) top 12", err);
      Matrix<float> Vf(n, 24), Df(24, 24, Location::kHOST);
      Vf.fill_random(9);
      st = si_eig<__half, float, float>(Ah, Vf, Df, 200, 1e-8f, false, 1e-12f);
      double errf = 0;
      for (int i = 0; i < 12; ++i) errf = std::fmax(errf, std::fabs((double)Df(i, i) - ex[i]) / ex[i]);
      check(st == MSVDStatus::kSuccess && errf < 1e-6, "si_eig<__half,float,float> top 12", errf);
      // ofrr_eig on half-precision projected matrices of an exact invariant basis
      Matrix<float> Dr(12, 12, Location::kHOST);
      Matrix<float> Vr = Vf.submatrix(0, 0, n, 12).cast<float>();
      st = rr_eig<__half, float, float>(Ah, Vr, Dr, false, 1e-10f);
      double err2 = 0;
      for (int i = 0; i < 6; ++i) err2 = std::fmax(err2, std::fabs((double)Dr(i, i) - ex[i]) / ex[i]);
      check(st == MSVDStatus::kSuccess && err2 < 1e-5, "rr_eig<__half,float,float> on a converged basis", err2);
      Matrix<__half> Gh(3, 3, Location::kHOST), Hh(3, 3, Location::kHOST);
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) {
            Gh(i, j) = __float2half(i == j ? 2.0f : 0.0f);
            Hh(i, j) = __float2half(i == j ? (float)(2 * (i + 1)) : 0.0f);
         }
      Matrix<float> Ve(3, 3, Location::kHOST), De(3, 3, Location::kHOST);
      st = ofrr_eig<__half, float, float>(Hh, Gh, Ve, De, 1e-10f);
      check(st == MSVDStatus::kSuccess && De(0, 0) == 3.0f && De(1, 1) == 2.0f && De(2, 2) == 1.0f,
            "ofrr_eig<__half,float,float> pencil (diag(2,4,6), 2 I)", De(0, 0));
   }
#endif
   // ---- special values
   {
      Matrix<double> A(8, 8, Location::kHOST);
      A(3, 4) = std::nan("");
      check(check_special_values(A) == MSVDStatus::kErrorNaN, "check_special_values NaN");
   }
   std::printf("%d failure(s)\n", g_fail);
   return g_fail ? 1 : 0;
}
