// capi_check.cpp - the library's C ABI driven the way a ported caller drives it: headers from
// include/, the shared library linked directly, device memory from hipMalloc, results checked
// against plain loops on the host.
//
//   capi_check host   entry points that do no device work (version, planner, knobs, argument
//                     validation)
//   capi_check gpu    the twins of include/cl_marshal.h (the reference's 3-stage route among
//                     them) and of include/marshal.h on a device buffer, each checked
//
// Prints "capi <mode> ok" and exits 0 when every check passes.
#include "hip_marshal.h"
#include "marshal.h"

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

static int g_fail = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "capi_check:%d: %s\n", __LINE__, #cond);   \
      g_fail++;                                                       \
    }                                                                 \
  } while (0)

static int host_checks() {
  CHECK(std::strstr(hip_marshal_version(), "gfx950") != nullptr);
  int64_t a = 0, b = 0;
  int stages = 0;
  CHECK(!hip_marshal_plan_transpose(4096, 4096, 8, &a, &b, &stages));
  CHECK(stages == 8);  // square: tile-pair swaps
  CHECK(hip_marshal_plan_transpose(0, 5, 4, &a, &b, &stages));  // invalid shape: error
  CHECK(hip_marshal_get_tuning("bs_upt") == 16);
  CHECK(hip_marshal_set_tuning("bs_upt", 0));                    // below its minimum
  CHECK(hip_marshal_set_tuning("no_such_key", 1));
  CHECK(hip_marshal_get_tuning("bs_upt") == 16);
  CHECK(hip_marshal_flag_bytes_0100(1, 16, 1 << 18, 64, 4) == size_t(16) * (1 << 18) / 8);
  // argument validation happens before any device call
  CHECK(hip_transpose_es(nullptr, nullptr, 4, 2, 3, 4, 5, 1, 3, nullptr));
  CHECK(hip_transpose_es(nullptr, reinterpret_cast<void *>(256), 3, 2, 3, 4, 5, 1, 3, nullptr));
  CHECK(hip_aos_asta_es(nullptr, reinterpret_cast<void *>(256), 4, 10, 3, 4, 0, nullptr));
  CHECK(hip_padding_es(nullptr, reinterpret_cast<void *>(256), 4, 9, 3, 8, nullptr));
  return g_fail;
}

template <typename T>
static std::vector<T> pattern(size_t n) {
  std::vector<T> v(n);
  for (size_t i = 0; i < n; i++) v[i] = T(i * 2654435761u % 1000003u);
  return v;
}

template <typename T>
static bool upload(T *d, const std::vector<T> &h) {
  return hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

template <typename T>
static std::vector<T> download(const T *d, size_t n) {
  std::vector<T> h(n);
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
    h.clear();
  return h;
}

static int gpu_checks() {
  // the reference's 3-stage route on a (10 * 18) x (10 * 72) int matrix, then back
  {
    const int A = 10, a = 18, B = 10, b = 72, M = A * a, N = B * b;
    const auto x = pattern<int>(size_t(M) * N);
    int *d = nullptr;
    CHECK(hipMalloc(&d, x.size() * sizeof(int)) == hipSuccess);
    CHECK(upload(d, x));
    uint64_t et = 0;
    CHECK(!hip_transpose(nullptr, d, A, a, B, b, 1, 3, &et));
    CHECK(et > 0);  // profiling is on by default: the call's time is added
    auto y = download(d, x.size());
    CHECK(y.size() == x.size());
    bool ok = true;
    for (int i = 0; i < M && ok && !y.empty(); i++)
      for (int j = 0; j < N; j++)
        if (y[size_t(j) * M + i] != x[size_t(i) * N + j]) {
          ok = false;
          break;
        }
    CHECK(ok);
    const uint64_t et1 = et;
    CHECK(!hip_transpose(nullptr, d, B, b, A, a, 1, 3, &et));  // N x M back to M x N
    CHECK(et > et1);                                            // accumulated, not replaced
    CHECK(download(d, x.size()) == x);
    CHECK(hipFree(d) == hipSuccess);
  }
  // AoS -> ASTA (LDS-staged, and by cycle following) and SoA -> ASTA, float, h = 4096 x w = 16,
  // tile 64: ASTA[i / t][j][i % t] = AoS[i][j] = SoA[j][i]
  {
    const int h = 4096, w = 16, t = 64;
    const auto x = pattern<float>(size_t(h) * w);
    std::vector<float> asta(x.size());
    for (int i = 0; i < h; i++)
      for (int j = 0; j < w; j++) asta[size_t(i / t) * w * t + size_t(j) * t + i % t] = x[size_t(i) * w + j];
    std::vector<float> soa(x.size());
    for (int i = 0; i < h; i++)
      for (int j = 0; j < w; j++) soa[size_t(j) * h + i] = x[size_t(i) * w + j];
    float *d = nullptr;
    CHECK(hipMalloc(&d, x.size() * sizeof(float)) == hipSuccess);
    CHECK(upload(d, x));
    CHECK(!hip_aos_asta_bs(nullptr, d, h, w, t));
    CHECK(download(d, x.size()) == asta);
    CHECK(upload(d, x));
    CHECK(!hip_aos_asta_pttwac(nullptr, d, h, w, t, 1, 1));
    CHECK(download(d, x.size()) == asta);
    CHECK(upload(d, x));
    CHECK(!hip_aos_asta(nullptr, d, h, w, t, 1));
    CHECK(download(d, x.size()) == asta);
    // include/marshal.h: the reference's CUDA-side names, null stream, timer ignored
    CHECK(upload(d, x));
    CHECK(!gpu_aos_asta_bs_float(d, h, w, t, nullptr));
    CHECK(download(d, x.size()) == asta);
    CHECK(upload(d, x));
    CHECK(!gpu_aos_asta_pttwac(d, h, w, t, nullptr));
    CHECK(download(d, x.size()) == asta);
    CHECK(upload(d, soa));
    CHECK(!gpu_soa_asta_pttwac(d, h, w, t, nullptr));
    CHECK(download(d, x.size()) == asta);
    CHECK(hipFree(d) == hipSuccess);
  }
  // double AoS -> ASTA through marshal.h
  {
    const int h = 1024, w = 24, t = 32;
    const auto x = pattern<double>(size_t(h) * w);
    std::vector<double> asta(x.size());
    for (int i = 0; i < h; i++)
      for (int j = 0; j < w; j++) asta[size_t(i / t) * w * t + size_t(j) * t + i % t] = x[size_t(i) * w + j];
    double *d = nullptr;
    CHECK(hipMalloc(&d, x.size() * sizeof(double)) == hipSuccess);
    CHECK(upload(d, x));
    CHECK(!gpu_aos_asta_bs_double(d, h, w, t, nullptr));
    CHECK(download(d, x.size()) == asta);
    CHECK(hipFree(d) == hipSuccess);
  }
  // padding [y][x] -> [y][pad] (new columns zeroed) and back
  {
    const int xs = 1001, ys = 300, pad = 1024;
    std::vector<int> x(size_t(ys) * pad, 0);
    const auto body = pattern<int>(size_t(ys) * xs);
    std::copy(body.begin(), body.end(), x.begin());
    std::vector<int> padded(x.size(), 0);
    for (int r = 0; r < ys; r++)
      for (int c = 0; c < xs; c++) padded[size_t(r) * pad + c] = body[size_t(r) * xs + c];
    int *d = nullptr;
    CHECK(hipMalloc(&d, x.size() * sizeof(int)) == hipSuccess);
    CHECK(upload(d, x));
    uint64_t et = 0;
    CHECK(!hip_padding(nullptr, d, xs, ys, pad, &et));
    CHECK(download(d, x.size()) == padded);
    CHECK(!hip_unpadding(nullptr, d, xs, ys, pad, &et));
    CHECK(download(d, x.size()) == x);
    CHECK(hipFree(d) == hipSuccess);
  }
  hip_marshal_finalize();
  return g_fail;
}

int main(int argc, char **argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  const int fails = gpu ? gpu_checks() : host_checks();
  if (fails) {
    std::printf("capi %s: %d checks failed\n", gpu ? "gpu" : "host", fails);
    return 1;
  }
  std::printf("capi %s ok\n", gpu ? "gpu" : "host");
  return 0;
}
