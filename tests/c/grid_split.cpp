// Host-only check of trid::for_each_grid_chunk (tridsolver_amd/csrc/
// trid_internal.h): every sub-batch's grid stays below 2^32 work-items, the
// sub-batches cover every system exactly once, in order, and a grid that
// fits is launched whole. Prints "ok" and exits 0 on success. Makes no HIP
// call, so it runs without a GPU.
#include <cstdio>
#include <cstdlib>

#include "../../tridsolver_amd/csrc/trid_internal.h"

using trid::Geom;

static int fail(const char *what, long long a, long long b) {
  std::printf("FAIL %s: %lld %lld\n", what, a, b);
  return 1;
}

// n0 x n1 systems, st0 = 1 and st1 = n0, so an element offset names the
// first system of a sub-batch.
static int check(long long n0, long long n1, int W, int T, long long want_calls) {
  Geom g;
  g.N = 2;
  g.stride = n0 * n1;
  g.n0 = n0;
  g.n1 = n1;
  g.st0 = 1;
  g.st1 = n0;
  long long calls = 0, next = 0, bad = 0;
  const hipError_t e = trid::for_each_grid_chunk(g, W, T, [&](const Geom &s, int64_t off) {
    ++calls;
    const long long blocks = (s.n0 + W - 1) / W * s.n1;
    if (blocks * T > (long long)UINT32_MAX) ++bad;
    // the sub-batch starts where the previous one ended (i0-major within
    // i1 when n0 is split, whole i1 rows otherwise)
    if (off != next) ++bad;
    if (s.n0 != n0 && s.n1 != 1) ++bad;
    if (s.n0 == n0) next = off + s.n0 * s.n1;
    else next = off + s.n0;
    if (s.N != g.N || s.stride != g.stride || s.st0 != g.st0 || s.st1 != g.st1) ++bad;
    return hipSuccess;
  });
  if (e != hipSuccess) return fail("status", (long long)e, 0);
  if (bad) return fail("sub-batch", bad, calls);
  if (next != n0 * n1) return fail("coverage", next, n0 * n1);
  if (want_calls > 0 && calls != want_calls) return fail("calls", calls, want_calls);
  return 0;
}

int main() {
  int r = 0;
  r |= check(512, 512, 16, 512, 1);                 // fits: one call
  r |= check(2147483647LL, 1, 128, 256, 0);         // 2^31-1 rows of length 2 along x
  r |= check(1000, 2200000, 16, 512, 0);            // split along i1
  r |= check(3000000000LL / 3, 3, 4, 1024, 0);      // i0 split in every i1 slice
  r |= check(1, 1, 1, 1024, 1);
  if (r) return 1;
  std::printf("ok\n");
  return 0;
}
