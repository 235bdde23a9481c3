// Host checks of the index arithmetic of flb_device.h / flb_internal.h:
// block_tile and block_tile_window with the scalar magic division give the
// tiles of the plain division form for every block index of many grid shapes,
// div_by equals n / d, band_reversed is a permutation in the order the odd
// steps walk, and plane_stride keeps planes 256-byte aligned, at least
// a plane long and 4 KiB past a multiple of 8 KiB apart.  Built and run by
// tests/test_block_mapping.py.
#include <cstdio>
#include <cstdint>
#include <vector>
#include "flb_internal.h"
using namespace flb;
// The reference form: plain divisions; in a full band of 32 rows block q of a
// column takes row (q mod 8) * 4 + q / 8 (XCD k = q mod 8 gets rows 4k..4k+3).
static uint32_t band_row(uint32_t q, uint32_t h) { return h == 32 ? (q % 8) * 4 + q / 8 : q; }
static void old_tile(uint32_t b, uint32_t tiles_x, uint32_t rows, uint32_t& tx, uint32_t& y) {
  const uint32_t per_band = kBandRows * tiles_x;
  const uint32_t band = b / per_band;
  const uint32_t r = b - band * per_band;
  const uint32_t y0 = band * kBandRows;
  const uint32_t h = rows - y0 < kBandRows ? rows - y0 : kBandRows;
  tx = r / h;
  y = y0 + band_row(r - tx * h, h);
}
static void old_window(const Params& p, uint32_t b, uint32_t& tx, uint32_t& y) {
  const uint32_t per_band = kBandRows * p.ncols;
  const uint32_t band = b / per_band;
  const uint32_t r = b - band * per_band;
  const uint32_t y0 = (p.band0 + band) * kBandRows;
  const uint32_t h = p.ny - y0 < kBandRows ? p.ny - y0 : kBandRows;
  const uint32_t c = r / h;
  tx = p.col0 + c;
  y = y0 + band_row(r - c * h, h);
}
int main() {
  uint64_t checked = 0; int bad = 0;
  const uint32_t txs[] = {1, 2, 3, 4, 5, 7, 8, 16, 31, 32, 33, 64, 100, 128, 255, 256, 1000, 4096};
  const uint32_t rs[] = {2, 3, 31, 32, 33, 63, 64, 65, 100, 128, 129, 256, 1000, 2048, 2049};
  for (uint32_t t : txs) for (uint32_t rows : rs) {
    const uint64_t nb = uint64_t(t) * rows;
    if (nb > (1u << 24)) continue;
    const uint32_t m = div_magic(uint64_t(kBandRows) * t);
    for (uint32_t b = 0; b < nb; ++b) {
      uint32_t a1, b1, a2, b2;
      old_tile(b, t, rows, a1, b1);
      block_tile(b, t, rows, m, div_magic(last_band_rows(rows)), a2, b2);
      if (a1 != a2 || b1 != b2) { if (bad++ < 5) printf("tile mismatch t=%u rows=%u b=%u\n", t, rows, b); }
      ++checked;
      // window: band0 1, col0 2, ncols t (a window inside a wider grid)
      Params p{}; p.ny = rows + 32; p.band0 = 1; p.col0 = 2; p.ncols = t; p.last_magic = div_magic(last_band_rows(p.ny));
      p.win_magic = div_magic(uint64_t(kBandRows) * t);
      uint32_t c1, y1, c2, y2;
      block_tile_window(p, b, c2, y2);
      old_window(p, b, c1, y1);
      if (c1 != c2 || y1 != y2) { if (bad++ < 5) printf("window mismatch t=%u rows=%u b=%u\n", t, rows, b); }
    }
  }
  // every tile of a grid is visited exactly once
  for (uint32_t t : {1u, 3u, 32u, 33u}) for (uint32_t rows : {31u, 32u, 64u, 100u}) {
    const uint32_t m = div_magic(uint64_t(kBandRows) * t), ml = div_magic(last_band_rows(rows));
    std::vector<int> seen(size_t(t) * rows, 0);
    for (uint32_t b = 0; b < t * rows; ++b) {
      uint32_t tx, y;
      block_tile(b, t, rows, m, ml, tx, y);
      if (tx >= t || y >= rows) { if (bad++ < 5) printf("out of grid t=%u rows=%u b=%u\n", t, rows, b); continue; }
      seen[size_t(y) * t + tx]++;
    }
    for (int v : seen) if (v != 1) { if (bad++ < 5) printf("tile not visited once t=%u rows=%u\n", t, rows); break; }
    ++checked;
  }
  // band_reversed (the odd steps' walk): a permutation of the launch's blocks,
  // the last band first, then the full bands from the second-to-last down, each
  // band's blocks in their forward order, and a block index congruent mod 8 to
  // the others of its band's rows (the XCD grouping of a full band is kept).
  for (uint32_t cols : {1u, 2u, 3u, 7u, 8u, 33u, 128u}) for (uint32_t rows : {1u, 5u, 31u, 32u, 33u, 64u, 100u, 2049u}) {
    Params p{};
    set_walk(p, rows, cols);
    const uint32_t per = kBandRows * cols, m = div_magic(per), n = rows * cols;
    std::vector<int> seen(n, 0);
    for (uint32_t b = 0; b < n; ++b) {
      const uint32_t f = band_reversed(b, p.walk_last, p.walk_bands, per, m);
      if (f >= n) { if (bad++ < 5) printf("reversed out of range cols=%u rows=%u b=%u\n", cols, rows, b); continue; }
      seen[f]++;
      const uint32_t band = f / per;
      const uint32_t want = b < p.walk_last ? p.walk_bands - 1 : p.walk_bands - 2 - (b - p.walk_last) / per;
      const uint32_t off = b < p.walk_last ? b : (b - p.walk_last) % per;
      if (band != want || f - band * per != off) { if (bad++ < 5) printf("reversed order cols=%u rows=%u b=%u\n", cols, rows, b); }
      if (band + 1 < p.walk_bands && (b - p.walk_last) % 8 != (f - band * per) % 8) { if (bad++ < 5) printf("reversed xcd cols=%u rows=%u b=%u\n", cols, rows, b); }
    }
    for (int v : seen) if (v != 1) { if (bad++ < 5) printf("reversed not a permutation cols=%u rows=%u\n", cols, rows); break; }
    ++checked;
  }
  // div_by against n / d on large numerators
  for (uint64_t d = 2; d < 200000; d += 37) {
    const uint32_t m = div_magic(d);
    for (uint32_t n : {0u, 1u, uint32_t(d - 1), uint32_t(d), uint32_t(d + 1), 0x7fffffffu, 0x7ffffffeu, 123456789u}) {
      if (div_by(n, uint32_t(d), m) != n / uint32_t(d)) { if (bad++ < 5) printf("div mismatch n=%u d=%llu\n", n, (unsigned long long)d); }
      ++checked;
    }
  }
  for (uint32_t elem : {2u, 4u, 8u}) {
    for (uint64_t cells : {64ull, 4096ull, 8192ull * 2048ull, 8192ull * 2048ull + 64ull,
                           1024ull * 256ull, 2112ull * 512ull, 123456ull * 64ull}) {
      const uint64_t np = plane_stride(cells, elem), bytes = np * elem;
      if (np < cells || bytes % 256 != 0 || bytes % 8192 != 4096 || bytes - cells * elem >= 12288) {
        if (bad++ < 5) printf("plane_stride cells=%llu elem=%u -> %llu\n", (unsigned long long)cells, elem, (unsigned long long)np);
      }
      ++checked;
    }
  }
  printf("checked %llu, mismatches %d\n", (unsigned long long)checked, bad);
  return bad != 0;
}
