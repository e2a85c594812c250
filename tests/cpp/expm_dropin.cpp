// A C++ caller written against the reference's calaman::expm interface
// (src/expm/interface.cppm usage block), built against include/calaman/expm.hpp
// and libcalaman_expm.so. Mirrors three reference cases
// (test/expm/expm_tests.cpp): exp(0) = I, exp(diag) = diag(exp), padded lde,
// plus exp(A) exp(-A) = I, for float, double, std::complex<float> and
// std::complex<double>; and the neighbouring templates in double and
// std::complex<double> -- matrix_norm1, lange (all four norms), lascl,
// paterson_stockmeyer and horner (test/lange/lange_tests.cpp,
// test/lascl/lascl_tests.cpp, test/paterson_stockmeyer/paterson_stockmeyer_tests.cpp,
// test/horner/horner_tests.cpp) -- against host arithmetic.
// Exit status 0 when every check passes.
#include <calaman/expm.hpp>

#include <cmath>
#include <complex>
#include <cstdio>
#include <limits>
#include <utility>
#include <vector>

template <typename T>
static int run_case(const char *what, int n, int lde, const std::vector<T> &a,
                    std::vector<T> &out, calaman::ExpmPlan *plan) {
  T *dA = nullptr, *dE = nullptr;
  int *dInfo = nullptr;
  void *dWork = nullptr;
  std::size_t bytes = 0;
  if (!calaman::expm_bufferSize<T>(nullptr, n, &bytes).ok()) return 1;
  if (hipMalloc(&dA, sizeof(T) * n * n) || hipMalloc(&dE, sizeof(T) * lde * n) ||
      hipMalloc(&dWork, bytes) || hipMalloc(&dInfo, 2 * sizeof(int)))
    return 1;
  (void)hipMemcpy(dA, a.data(), sizeof(T) * n * n, hipMemcpyHostToDevice);
  (void)hipMemcpy(dE, out.data(), sizeof(T) * lde * n, hipMemcpyHostToDevice);
  hipStream_t stream;
  if (hipStreamCreate(&stream) != hipSuccess) return 1;
  const calaman::Status st =
      calaman::expm<T>(nullptr, nullptr, stream, n, dA, n, dE, lde, dWork, bytes, dInfo, plan);
  (void)hipStreamSynchronize(stream);
  (void)hipMemcpy(out.data(), dE, sizeof(T) * lde * n, hipMemcpyDeviceToHost);
  (void)hipStreamDestroy(stream);
  (void)hipFree(dA);
  (void)hipFree(dE);
  (void)hipFree(dWork);
  (void)hipFree(dInfo);
  if (!st.ok()) std::printf("%s: %s\n", what, st.name());
  return st.ok() ? 0 : 1;
}

template <typename T>
static int check_type(const char *tname) {
  int fails = 0;
  using R = calaman::RealType<T>;
  const int n = 6, lde = 9;
  const double eps = std::numeric_limits<R>::epsilon();
  // exp(0) = I into a padded buffer whose padding must survive
  std::vector<T> zero(n * n, T(0)), out(lde * n, T(-7.0));
  calaman::ExpmPlan plan;
  fails += run_case<T>("zero", n, lde, zero, out, &plan);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < lde; ++i) {
      const T want = (i >= n) ? T(-7.0) : (i == j ? T(1.0) : T(0.0));
      if (out[i + j * lde] != want) {
        ++fails;
        std::printf("%s exp(0) (%d,%d) wrong\n", tname, i, j);
      }
    }
  // exp(diag) = diag(exp)
  std::vector<T> d(n * n, T(0));
  for (int i = 0; i < n; ++i) d[i + i * n] = T(0.3 * i - 0.7);
  std::vector<T> e(n * n, T(0));
  fails += run_case<T>("diag", n, n, d, e, &plan);
  for (int i = 0; i < n; ++i)
    if (std::abs(e[i + i * n] - std::exp(d[i + i * n])) > 64 * eps * n) {
      ++fails;
      std::printf("%s exp(diag) %d wrong\n", tname, i);
    }
  // exp(A) exp(-A) = I for a dense A with ||A||_1 ~ 3
  std::vector<T> a(n * n), na(n * n), e1(n * n), e2(n * n);
  for (int k = 0; k < n * n; ++k) {
    a[k] = T(std::sin(1.0 + k) * 0.5);
    if constexpr (!std::is_same_v<T, R>) a[k] += T(R(0), R(std::cos(2.0 * k) * 0.3));
    na[k] = -a[k];
  }
  fails += run_case<T>("A", n, n, a, e1, &plan);
  fails += run_case<T>("-A", n, n, na, e2, &plan);
  double err = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      T s = 0;
      for (int k = 0; k < n; ++k) s += e1[i + k * n] * e2[k + j * n];
      err = std::fmax(err, static_cast<double>(std::abs(s - T(i == j ? 1.0 : 0.0))));
    }
  // about 1e-13 in double and 1e-5 in float (||exp(A)|| ||exp(-A)|| ~ e^6)
  if (err > 4096 * eps) {
    ++fails;
    std::printf("%s exp(A)exp(-A) err %.3e\n", tname, err);
  }
  std::printf("%s: plan m=%d s=%d gemms=%d, inverse err %.2e, %s\n", tname, plan.m, plan.s,
              plan.num_gemms, err, fails ? "FAIL" : "ok");
  return fails;
}

// host power series sum_k c_k A^k (column-major n x n)
template <typename T>
static std::vector<T> host_poly(int n, const std::vector<T> &a, const std::vector<T> &c) {
  std::vector<T> res(n * n, T(0)), pw(n * n, T(0)), nxt(n * n);
  for (int i = 0; i < n; ++i) pw[i + i * n] = T(1);
  for (std::size_t k = 0; k < c.size(); ++k) {
    for (int e = 0; e < n * n; ++e) res[e] += c[k] * pw[e];
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        T s = 0;
        for (int l = 0; l < n; ++l) s += pw[i + l * n] * a[l + j * n];
        nxt[i + j * n] = s;
      }
    pw.swap(nxt);
  }
  return res;
}

template <typename T>
static double max_abs_diff(const std::vector<T> &x, const std::vector<T> &y) {
  double d = 0;
  for (std::size_t e = 0; e < x.size(); ++e) d = std::fmax(d, std::abs(x[e] - y[e]));
  return d;
}

template <typename T>
static int check_neighbours(const char *tname) {
  using R = calaman::RealType<T>;
  int fails = 0;
  const int m = 7, n = 5, lda = 9;
  std::vector<T> a(lda * n, T(-9.0));  // padding rows must not count
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      a[i + j * lda] = T(std::sin(1.0 + i + 3.0 * j));
      if constexpr (!std::is_same_v<T, R>) a[i + j * lda] += T(R(0), R(std::cos(0.5 * i - j)));
    }
  T *dA = nullptr;
  R *dR = nullptr, *dColsum = nullptr;
  if (hipMalloc(&dA, sizeof(T) * lda * n) || hipMalloc(&dR, sizeof(R)) ||
      hipMalloc(&dColsum, sizeof(R) * 64))
    return 1;
  (void)hipMemcpy(dA, a.data(), sizeof(T) * lda * n, hipMemcpyHostToDevice);
  hipStream_t stream;
  if (hipStreamCreate(&stream) != hipSuccess) return 1;
  // lange: the four norms of the m x n block
  double one = 0, inf = 0, mx = 0, fro = 0;
  for (int j = 0; j < n; ++j) {
    double cs = 0;
    for (int i = 0; i < m; ++i) {
      const double v = std::abs(a[i + j * lda]);
      cs += v;
      mx = std::fmax(mx, v);
      fro += v * v;
    }
    one = std::fmax(one, cs);
  }
  for (int i = 0; i < m; ++i) {
    double rs = 0;
    for (int j = 0; j < n; ++j) rs += std::abs(a[i + j * lda]);
    inf = std::fmax(inf, rs);
  }
  fro = std::sqrt(fro);
  const std::pair<calaman::MatrixNorm, double> norms[4] = {{calaman::MatrixNorm::one, one},
                                                           {calaman::MatrixNorm::inf, inf},
                                                           {calaman::MatrixNorm::max_abs, mx},
                                                           {calaman::MatrixNorm::frobenius, fro}};
  for (const auto &[which, want] : norms) {
    R got = -1;
    if (!calaman::lange<T>(stream, which, m, n, dA, lda, dR).ok()) ++fails;
    (void)hipMemcpyAsync(&got, dR, sizeof(R), hipMemcpyDeviceToHost, stream);
    (void)hipStreamSynchronize(stream);
    if (std::fabs(got - want) > 1e-13 * want) {
      ++fails;
      std::printf("%s lange(%d) %.17g want %.17g\n", tname, static_cast<int>(which), double(got),
                  want);
    }
  }
  // matrix_norm1 of the leading n x n block (true modulus for complex)
  double sq1 = 0;
  for (int j = 0; j < n; ++j) {
    double cs = 0;
    for (int i = 0; i < n; ++i) cs += std::abs(a[i + j * lda]);
    sq1 = std::fmax(sq1, cs);
  }
  const R nrm = calaman::matrix_norm1<T>(stream, n, dA, lda, dColsum);
  if (std::fabs(nrm - sq1) > 1e-13 * sq1) {
    ++fails;
    std::printf("%s matrix_norm1 %.17g want %.17g\n", tname, double(nrm), sq1);
  }
  // lascl by a power of two: exact, padding untouched
  if (!calaman::lascl<T>(stream, R(4), R(1), m, n, dA, lda).ok()) ++fails;
  std::vector<T> back(lda * n);
  (void)hipStreamSynchronize(stream);
  (void)hipMemcpy(back.data(), dA, sizeof(T) * lda * n, hipMemcpyDeviceToHost);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < lda; ++i) {
      const T want = (i < m) ? a[i + j * lda] * T(0.25) : T(-9.0);
      if (back[i + j * lda] != want) {
        ++fails;
        std::printf("%s lascl (%d,%d)\n", tname, i, j);
      }
    }
  // Paterson-Stockmeyer and Horner on a context bound to the stream: every block
  // size against the host series, PS == Horner, the zero polynomial exactly 0
  const int k = 6, degree = 7;
  std::vector<T> sa(k * k), coef(degree + 1);
  for (int e = 0; e < k * k; ++e) sa[e] = T(0.3 * std::sin(0.7 * e));
  for (int c = 0; c <= degree; ++c) {
    coef[c] = T(1.0 / (1.0 + c));
    if constexpr (!std::is_same_v<T, R>) coef[c] += T(R(0), R(0.1 * c));
  }
  const std::vector<T> want = host_poly(k, sa, coef);
  clm_ctx *ctx = nullptr;
  if (clm_create(&ctx) != CLM_STATUS_SUCCESS || clm_set_stream(ctx, stream) != CLM_STATUS_SUCCESS)
    return 1;
  T *dS = nullptr, *dP = nullptr, *dC = nullptr;
  void *dW = nullptr;
  const std::size_t wbytes = calaman::paterson_stockmeyer_bufferSize<T>(k, degree, 1) +
                             calaman::horner_bufferSize<T>(k) + (1 << 16);
  if (hipMalloc(&dS, sizeof(T) * k * k) || hipMalloc(&dP, sizeof(T) * k * k) ||
      hipMalloc(&dC, sizeof(T) * (degree + 1)) || hipMalloc(&dW, wbytes))
    return 1;
  (void)hipMemcpy(dS, sa.data(), sizeof(T) * k * k, hipMemcpyHostToDevice);
  (void)hipMemcpy(dC, coef.data(), sizeof(T) * (degree + 1), hipMemcpyHostToDevice);
  std::vector<T> got(k * k), horner_out(k * k);
  for (int s = 0; s <= degree + 1; ++s) {
    if (!calaman::paterson_stockmeyer<T>(ctx, k, dC, degree, dS, k, dP, k, dW, wbytes, s).ok())
      ++fails;
    (void)hipStreamSynchronize(stream);
    (void)hipMemcpy(got.data(), dP, sizeof(T) * k * k, hipMemcpyDeviceToHost);
    if (max_abs_diff(got, want) > 1e-13) {
      ++fails;
      std::printf("%s paterson_stockmeyer s=%d err %.3e\n", tname, s, max_abs_diff(got, want));
    }
  }
  if (!calaman::horner<T>(ctx, k, dC, degree, dS, k, dP, k, dW, wbytes).ok()) ++fails;
  (void)hipStreamSynchronize(stream);
  (void)hipMemcpy(horner_out.data(), dP, sizeof(T) * k * k, hipMemcpyDeviceToHost);
  if (max_abs_diff(horner_out, want) > 1e-13 || max_abs_diff(horner_out, got) > 1e-13) {
    ++fails;
    std::printf("%s horner err %.3e\n", tname, max_abs_diff(horner_out, want));
  }
  std::vector<T> zeros(degree + 1, T(0));
  (void)hipMemcpy(dC, zeros.data(), sizeof(T) * (degree + 1), hipMemcpyHostToDevice);
  if (!calaman::paterson_stockmeyer<T>(ctx, k, dC, degree, dS, k, dP, k, dW, wbytes).ok()) ++fails;
  (void)hipStreamSynchronize(stream);
  (void)hipMemcpy(got.data(), dP, sizeof(T) * k * k, hipMemcpyDeviceToHost);
  for (const T &v : got)
    if (v != T(0)) {
      ++fails;
      std::printf("%s zero polynomial not exactly 0\n", tname);
      break;
    }
  (void)clm_destroy(ctx);
  (void)hipStreamDestroy(stream);
  for (void *ptr : {static_cast<void *>(dA), static_cast<void *>(dR), static_cast<void *>(dColsum),
                    static_cast<void *>(dS), static_cast<void *>(dP), static_cast<void *>(dC), dW})
    (void)hipFree(ptr);
  std::printf("%s neighbours (lange, matrix_norm1, lascl, paterson_stockmeyer, horner): %s\n",
              tname, fails ? "FAIL" : "ok");
  return fails;
}

int main() {
  int fails = check_type<double>("double") + check_type<std::complex<double>>("complex<double>") +
              check_type<float>("float") + check_type<std::complex<float>>("complex<float>") +
              check_neighbours<double>("double") +
              check_neighbours<std::complex<double>>("complex<double>");
  std::printf("%s\n", fails ? "FAILED" : "PASSED");
  return fails ? 1 : 0;
}
