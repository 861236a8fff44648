// fftahip_percall: the per-call path timed from native code.  One
// fftahip_execute (mul!(y, p, x), src/plan.jl:157-167) per transform on one
// stream, plans built beforehand (one per distinct length and direction),
// arrays resident in device memory; no interpreter between the calls.
//
//   fftahip_percall <calls file> <precision 32|64> [repeats]
//
// The calls file holds int64 count, then count int64 lengths, then count
// int32 directions (-1 / +1).  Prints one JSON line: transforms, distinct
// plans, device time of the timed loop (events on the stream) and host wall
// time of the same loop (issue included), each as transforms per second.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <utility>
#include <vector>

#include "fftahip.h"

static int die(const char* what, fftahip_status st) {
  std::fprintf(stderr, "fftahip_percall: %s: %s (%s)\n", what, fftahip_status_string(st), fftahip_last_error());
  return 1;
}

static int hip_die(const char* what, hipError_t e) {
  std::fprintf(stderr, "fftahip_percall: %s: %s\n", what, hipGetErrorString(e));
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <calls file> <precision 32|64> [repeats]\n", argv[0]);
    return 2;
  }
  const int prec = std::atoi(argv[2]);
  const int repeats = argc > 3 ? std::max(1, std::atoi(argv[3])) : 1;
  if (prec != 32 && prec != 64) {
    std::fprintf(stderr, "precision must be 32 or 64\n");
    return 2;
  }
  std::FILE* f = std::fopen(argv[1], "rb");
  if (!f) {
    std::perror(argv[1]);
    return 2;
  }
  int64_t count = 0;
  if (std::fread(&count, sizeof(count), 1, f) != 1 || count <= 0 || count > (int64_t(1) << 24)) {
    std::fprintf(stderr, "bad calls file\n");
    std::fclose(f);
    return 2;
  }
  std::vector<int64_t> len((size_t)count);
  std::vector<int32_t> dir((size_t)count);
  const bool ok = std::fread(len.data(), sizeof(int64_t), (size_t)count, f) == (size_t)count &&
                  std::fread(dir.data(), sizeof(int32_t), (size_t)count, f) == (size_t)count;
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "short calls file\n");
    return 2;
  }
  std::vector<int64_t> off((size_t)count);
  int64_t total = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (len[i] < 1 || (dir[i] != -1 && dir[i] != 1)) {
      std::fprintf(stderr, "bad call %lld\n", (long long)i);
      return 2;
    }
    off[i] = total;
    total += len[i];
  }
  const size_t esz = prec == 32 ? 8 : 16;

  // one plan per distinct (length, direction)
  std::map<std::pair<int64_t, int>, fftahip_plan> plans;
  fftahip_status st;
  for (int64_t i = 0; i < count; ++i) {
    const auto key = std::make_pair(len[i], (int)dir[i]);
    if (plans.count(key)) continue;
    fftahip_plan p = nullptr;
    const int64_t region[1] = {1};
    if ((st = fftahip_plan_create(&p, prec, dir[i], 1, &len[i], 1, region, 73))) return die("plan_create", st);
    plans[key] = p;
  }

  // inputs: a fixed pseudo-random sequence (xorshift), resident on the device
  std::vector<double> host((size_t)total * 2);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (auto& v : host) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    v = (double)(s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
  void *x = nullptr, *y = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&x, (size_t)total * esz)) != hipSuccess) return hip_die("hipMalloc", e);
  if ((e = hipMalloc(&y, (size_t)total * esz)) != hipSuccess) return hip_die("hipMalloc", e);
  if (prec == 64) {
    e = hipMemcpy(x, host.data(), (size_t)total * esz, hipMemcpyHostToDevice);
  } else {
    std::vector<float> h32(host.begin(), host.end());
    e = hipMemcpy(x, h32.data(), (size_t)total * esz, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) return hip_die("hipMemcpy", e);
  hipStream_t q;
  if ((e = hipStreamCreateWithFlags(&q, hipStreamNonBlocking)) != hipSuccess) return hip_die("hipStreamCreate", e);

  auto run = [&]() -> fftahip_status {
    for (int64_t i = 0; i < count; ++i) {
      fftahip_plan p = plans[std::make_pair(len[i], (int)dir[i])];
      fftahip_status r = fftahip_execute(p, 1, &len[i], (const char*)x + off[i] * esz, (char*)y + off[i] * esz, q);
      if (r) return r;
    }
    return FFTAHIP_SUCCESS;
  };
  // warm-up: uploads each shape's schedule, first launches
  if ((st = run())) return die("execute", st);
  if ((e = hipStreamSynchronize(q)) != hipSuccess) return hip_die("sync", e);

  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const auto t0 = std::chrono::steady_clock::now();
  (void)hipEventRecord(e0, q);
  for (int r = 0; r < repeats; ++r)
    if ((st = run())) return die("execute", st);
  (void)hipEventRecord(e1, q);
  if ((e = hipEventSynchronize(e1)) != hipSuccess) return hip_die("sync", e);
  const auto t1 = std::chrono::steady_clock::now();
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double wall_s = std::chrono::duration<double>(t1 - t0).count();
  const double n = (double)count * repeats;
  std::printf(
      "{\"transforms\": %lld, \"repeats\": %d, \"distinct_plans\": %zu, \"device_ms\": %.4f, \"wall_ms\": %.4f, "
      "\"transforms_per_s\": %.1f, \"wall_transforms_per_s\": %.1f}\n",
      (long long)count, repeats, plans.size(), (double)ms, wall_s * 1e3, n / (ms * 1e-3), n / wall_s);
  for (auto& kv : plans) fftahip_plan_destroy(kv.second);
  (void)hipFree(x);
  (void)hipFree(y);
  (void)hipStreamDestroy(q);
  return 0;
}
