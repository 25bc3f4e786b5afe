// plan.h -- compile-time Stockham FFT plans, shared by host and device code.
//
// A transform of length N is computed by TPT = N/P threads, each holding P
// points in registers.  Register slot i of thread t always holds position
// t + i*TPT at the start and at the end of a transform, so the first stage
// reads and the last stage writes straight from/to global memory (coalesced),
// and only the intermediate stages exchange data through LDS.
//
// Stage s has radix r[s] and span L[s] (product of the earlier radices).
// Butterfly j (0 <= j < N/r) reads positions j + m*N/r, multiplies input m by
// W_{L r}^{(j mod L) m}, runs a size-r DFT and writes position
// (j div L)*L*r + (j mod L) + m*L (the Stockham autosort ordering).
#pragma once

namespace nadia {

constexpr int kMaxStages = 12;

struct PlanDesc {
  int n, p, tpt, nst;
  int r[kMaxStages];
  int L[kMaxStages];
  int tw_off[kMaxStages + 1];  // offsets (in complex elements) of each stage's twiddles
};

// Radices are taken from {16, 8, 4, 3, 2}; each must divide both the
// remaining length and P (every thread runs P/r butterflies of radix r per
// stage).
constexpr PlanDesc make_plan(int n, int p) {
  PlanDesc d{};
  d.n = n;
  d.p = p;
  d.tpt = n / p;
  int m = n, L = 1, s = 0, off = 0;
  while (m > 1 && s < kMaxStages) {
    int r = 0;
    const int cand[5] = {16, 8, 4, 3, 2};
    for (int c = 0; c < 5; ++c)
      if (r == 0 && m % cand[c] == 0 && p % cand[c] == 0) r = cand[c];
    if (r == 0) break;  // unsupported length: caught by plan_ok
    d.r[s] = r;
    d.L[s] = L;
    d.tw_off[s] = off;
    off += L * (r - 1);
    L *= r;
    m /= r;
    ++s;
  }
  d.nst = (m == 1) ? s : -1;
  d.tw_off[s < kMaxStages ? s : kMaxStages] = off;
  return d;
}

constexpr int plan_twiddles(const PlanDesc& d) { return d.nst > 0 ? d.tw_off[d.nst] : 0; }

// Lengths with compiled single-launch passes: 2^k (8..4096) and 3*2^k
// (24..3072).  Other lengths take the any-length transforms (kern_any.inc)
// and the unfused iteration.
constexpr bool size_native(int n) {
  return n == 8 || n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 ||
         n == 1024 || n == 2048 || n == 4096 || n == 24 || n == 48 || n == 96 ||
         n == 192 || n == 384 || n == 768 || n == 1536 || n == 3072;
}

// Points per thread: row passes (contiguous transforms).  3*2^k lengths take
// 12 points (radix-4 and radix-3 stages): 24 points left the 384 and 768
// passes at two waves per SIMD (114-150 registers, 55 KB of LDS per
// workgroup) and 2-3x the time per pixel of the 2^k lengths.  Rows of 256,
// 512 and 1024 take 4 points (radix-4 stages, twice the threads per row):
// their one-job passes are latency-bound, and the row pass of one 256^2 /
// 512^2 / 1024^2 job gets 20% / 18% / 15% faster (batches unchanged); at
// 2048 the row pass is bandwidth-bound and 4 points cost 16%.
constexpr int row_points(int n) {
  if (n % 3 == 0) return n < 12 ? n : 12;
  if (n == 256 || n == 512 || n == 1024) return 4;
  return n < 8 ? n : 8;
}

// Points per thread for column passes: 16 for 2^k >= 2048 (radix-16
// stages, one LDS exchange fewer than the rows' radix-8 plan, and half the
// threads per column, so twice the columns per workgroup: 2048^2 FP32
// column pass 28.4 -> 26.5 us); 8 for 2^k below that (at 1024 radix-16
// columns measured 4% slower, 4 points 2% slower); 3*2^k as the rows.
constexpr int col_points(int n) {
  if (n % 3 == 0) return row_points(n);
  if (n >= 2048) return 16;
  return n < 8 ? n : 8;
}

// LDS pad shift of the exchanges (one pad element per 2^shift positions):
// column tiles of 16-point plans use 16, other column tiles 8; rows use 16
// (against 8: FP64 64-job batch 79.5k -> 82.3k job-iterations/s, FP32
// unchanged; one pad per 4 positions made the 2048^2 row pass 22% slower).
constexpr int col_pad_shift(int n) { return col_points(n) == 16 ? 4 : 3; }
constexpr int row_pad_shift(int) { return 4; }

// Transforms (rows) per row-pass workgroup: about 256 threads.
constexpr int row_batch(int n) {
  int tpt = n / row_points(n);
  int b = 256 / tpt;
  return b < 1 ? 1 : b;
}

// Columns per column-pass workgroup for element size `esz` (8 = fp32 complex,
// 16 = fp64 complex): up to 64-byte row segments, LDS tile N*C*esz <= 128 KiB,
// at most 1024 threads (512 in fp64, whose projection epilogues need more
// than the 128 registers a 1024-thread workgroup leaves each thread).
constexpr int col_batch(int n, int esz) {
  int c = 64 / esz;
  const int maxthr = esz > 8 ? 512 : 1024;
  while (c > 1 && (n * c * esz > 128 * 1024 || c * (n / col_points(n)) > maxthr)) c /= 2;
  return c;
}

// ---- tile variants chosen at run time by the autotuner --------------------
// Every variant of a length keeps the radix plan and the points per thread,
// so all of them produce bit-identical results; they differ in transforms per
// workgroup and in the occupancy target given to the register allocator.
struct TileCfg {
  int p;        // points per thread
  int batch;    // transforms (columns or rows) per workgroup at a time
  int threads;  // workgroup size
  int minw;     // waves per SIMD the register allocator must allow
};

constexpr int kColVariants = 3;
constexpr int kRowVariants = 2;

// lengths that get more than the default variant: 256..4096, 2^k and 3*2^k
constexpr bool tuned_size(int n) {
  return n == 256 || n == 384 || n == 512 || n == 768 || n == 1024 || n == 1536 || n == 2048 || n == 3072 ||
         n == 4096;
}

// Column variants: 0 default; 1 registers limited so that two workgroups
// share a CU; 2 half the columns per workgroup.
constexpr TileCfg col_cfg(int n, int esz, int v) {
  TileCfg c{col_points(n), col_batch(n, esz), 0, 1};
  if (!tuned_size(n)) v = 0;
  if (v == 2) c.batch = c.batch > 1 ? c.batch / 2 : 1;  // smaller workgroups, more per CU
  c.threads = c.batch * (n / c.p);
  if (v == 1) {  // ask for registers that let two workgroups share a CU
    const int w = c.threads / 128;
    if (n * c.batch * esz * 2 <= 160 * 1024 && w >= 2) c.minw = w > 8 ? 8 : w;
  }
  return c;
}

// Width of the column groups the error partials are summed over: the
// narrowest column tile among the variants of the length, so that every
// variant's tile holds whole groups and the partials do not depend on it.
constexpr int col_group(int n, int esz) { return col_cfg(n, esz, 2).batch; }

// Row variants: 0 default; 1 the same workgroup with registers limited so
// that 8 waves (fp32; 4 in fp64) share a SIMD.  esz: bytes per complex
// element.
constexpr TileCfg row_cfg(int n, int v, int esz = 8) {
  TileCfg c{row_points(n), row_batch(n), 0, 1};
  if (!tuned_size(n)) v = 0;
  if (v == 1) c.minw = esz > 8 ? 4 : 8;
  c.threads = c.batch * (n / c.p);
  return c;
}

}  // namespace nadia
