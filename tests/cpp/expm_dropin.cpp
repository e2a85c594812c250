// A C++ caller written against the reference's calaman::expm interface
// (src/expm/interface.cppm usage block), built against include/calaman/expm.hpp
// and libcalaman_expm.so. Mirrors three reference cases
// (test/expm/expm_tests.cpp): exp(0) = I, exp(diag) = diag(exp), padded lde,
// plus exp(A) exp(-A) = I, for float, double, std::complex<float> and
// std::complex<double>.
// Exit status 0 when every check passes.
#include <calaman/expm.hpp>

#include <cmath>
#include <complex>
#include <cstdio>
#include <limits>
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

int main() {
  int fails = check_type<double>("double") + check_type<std::complex<double>>("complex<double>") +
              check_type<float>("float") + check_type<std::complex<float>>("complex<float>");
  std::printf("%s\n", fails ? "FAILED" : "PASSED");
  return fails ? 1 : 0;
}
