// fft_device.h -- device building blocks: complex helpers, radix-2/3/4/8
// butterflies and the register/LDS Stockham stage loop (see plan.h).
#pragma once
#include <hip/hip_runtime.h>

#include "plan.h"

namespace nadia {

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };
template <typename T> using c2_t = typename Vec2<T>::type;

// ---- buffer-resource memory access ----------------------------------------
// Tile loads and stores go through buffer instructions: one per-thread 32-bit
// byte offset (VGPR) shared by all P slots of a thread, the slot stride in the
// scalar offset (SGPR), so the P strided accesses of a column tile cost no
// per-slot address registers.  Raw buffers with stride 0: num_records is the
// size in bytes; accesses beyond it read 0 and write nothing.
typedef unsigned int nadia_u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
  // 0x00020000: the gfx9-family descriptor word 3 (32-bit data, raw access)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

template <typename C>
__device__ __forceinline__ C bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(C) == 8) return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// 16-byte elements are stored as two 8-byte halves.  A buffer store of more
// than 64 bits of data needs a wait state before a VALU instruction may
// overwrite its data registers; the compiler does not insert it when the
// store's soffset is an SGPR, and the FP64 column pass then wrote a few
// lanes' values after their low dword had been replaced (nondeterministic
// errors of ~1e-9 in one slot of some threads).  64-bit stores are not
// subject to it.  The second half goes through soffset + 8 so that the two
// stores are not merged back into one 128-bit store.
template <typename C>
__device__ __forceinline__ void bstore(C v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(C) == 8) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(nadia_u32x2, v), r, voff, soff, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(nadia_u32x2, v.x), r, voff, soff, 0);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(nadia_u32x2, v.y), r, voff, soff + 8u, 0);
  }
}

template <typename T>
__device__ __forceinline__ T bload_real(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
// complex product with two explicit fused multiply-adds (the kernels are
// built without implicit contraction, so every instantiation rounds alike)
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return C{__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return float2{__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}
// a * (DIR * i)
template <int DIR, typename C> __device__ __forceinline__ C cmul_i(C a) {
  return DIR < 0 ? C{a.y, -a.x} : C{-a.y, a.x};
}
template <typename C, typename T> __device__ __forceinline__ C cscale(C a, T s) { return C{a.x * s, a.y * s}; }

// Size-R DFT in place, exponent sign DIR (-1 forward, +1 backward).
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(c2_t<T> (&a)[R]) {
  using C = c2_t<T>;
  if constexpr (R == 2) {
    C t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  } else if constexpr (R == 3) {
    const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
    C s = cadd(a[1], a[2]);
    C d = csub(a[1], a[2]);
    C m = C{a[0].x - (T)0.5 * s.x, a[0].y - (T)0.5 * s.y};
    a[0] = cadd(a[0], s);
    C u = cmul_i<DIR>(d);
    u = cscale(u, h);
    a[1] = cadd(m, u);
    a[2] = csub(m, u);
  } else if constexpr (R == 4) {
    C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    C t2 = cadd(a[1], a[3]), t3 = csub(a[1], a[3]);
    C u = cmul_i<DIR>(t3);
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, u);
    a[3] = csub(t1, u);
  } else if constexpr (R == 16) {
    // 4 x 4: radix-4 DFTs over n1 for each n2 (input n = 4 n1 + n2), the
    // twiddles W16^(n2 k1), radix-4 DFTs over n2; output k = k1 + 4 k2
    const T h = (T)0.70710678118654752440084436210485;  // cos(pi/4)
    const T c1 = (T)0.92387953251128675612818318939679;  // cos(pi/8)
    const T s1 = (T)0.38268343236508977172845998403040;  // sin(pi/8)
    C y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      C b[4] = {a[n2], a[n2 + 4], a[n2 + 8], a[n2 + 12]};
      dft<4, DIR, T>(b);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = b[k1];
    }
    // W16^j = cos(2 pi j / 16) + DIR i sin(2 pi j / 16)
    const T sd = DIR < 0 ? (T)-1 : (T)1;
    y[1][1] = cmul(y[1][1], C{c1, sd * s1});   // j = 1
    y[1][2] = cmul(y[1][2], C{h, sd * h});     // j = 2
    y[1][3] = cmul(y[1][3], C{s1, sd * c1});   // j = 3
    y[2][1] = cmul(y[2][1], C{h, sd * h});     // j = 2
    y[2][2] = cmul_i<DIR>(y[2][2]);            // j = 4
    y[2][3] = cmul(y[2][3], C{-h, sd * h});    // j = 6
    y[3][1] = cmul(y[3][1], C{s1, sd * c1});   // j = 3
    y[3][2] = cmul(y[3][2], C{-h, sd * h});    // j = 6
    y[3][3] = cmul(y[3][3], C{-c1, -sd * s1}); // j = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      C b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
      dft<4, DIR, T>(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = b[k2];
    }
  } else if constexpr (R == 8) {
    const T h = (T)0.70710678118654752440084436210485;  // sqrt(2)/2
    // even / odd radix-4 halves
    C e[4] = {a[0], a[2], a[4], a[6]};
    C o[4] = {a[1], a[3], a[5], a[7]};
    dft<4, DIR, T>(e);
    dft<4, DIR, T>(o);
    // o[k] *= W8^k, W8 = exp(DIR * i * pi / 4)
    C o1 = cmul_i<DIR>(o[1]);          // o1 * (DIR i)
    o1 = C{(o[1].x + o1.x) * h, (o[1].y + o1.y) * h};  // (1 + DIR i) h
    C o2 = cmul_i<DIR>(o[2]);
    C o3i = cmul_i<DIR>(o[3]);
    C o3 = C{(o3i.x - o[3].x) * h, (o3i.y - o[3].y) * h};  // (-1 + DIR i) h
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
  }
}

// LDS position of transform position p: one pad element after every 2^PS
// positions, so the strided writes of the first exchanges spread over the
// banks.  PS = 3 suits one element per position (row passes); a column tile
// stores CB elements per position, and for its 16-point plans one pad per
// 16 positions leaves no bank conflict in any exchange.
template <int N, int PS = 3> constexpr int lds_len() { return N + (N >> PS); }
template <int PS = 3> __device__ __forceinline__ int lpos(int p) { return p + (p >> PS); }

// Twiddles of stage S for the NB = P/R butterflies of one thread:
// w[q][m-1] = W_{L R}^{k m}, k = (t + q*TPT) mod L (sign -1 values).
template <typename T, int N, int P, int S>
struct StageTw {
  static constexpr PlanDesc pd = make_plan(N, P);
  static constexpr int R = (S < pd.nst) ? pd.r[S] : 2;
  c2_t<T> w[P / R][R - 1];
};

template <typename T, int N, int P, int S>
__device__ __forceinline__ void load_tw(StageTw<T, N, P, S>& w, int t, const c2_t<T>* __restrict__ tw) {
  constexpr PlanDesc pd = make_plan(N, P);
  if constexpr (S > 0 && S < pd.nst) {
    constexpr int R = pd.r[S], L = pd.L[S], TPT = N / P, NB = P / R;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int k = (t + q * TPT) % L;
      const c2_t<T>* p = tw + pd.tw_off[S] + k * (R - 1);
#pragma unroll
      for (int m = 1; m < R; ++m) w.w[q][m - 1] = p[m - 1];
    }
  }
}

// One transform of length N held by TPT = N/P threads, P points each.
// `lds` addresses position pos of this transform at lds[lpos(pos) * lstride]
// (the caller offsets the pointer per transform); `tw` is the plan's twiddle
// table (sign -1 values; DIR > 0 uses their conjugates).  The twiddles of
// stage S+1 are loaded before stage S's exchange barrier, so their latency
// overlaps the exchange.  Every thread of the workgroup must call this (it
// contains barriers).
template <typename T, int N, int P, int DIR, int S, int PS = 3>
__device__ __forceinline__ void fft_stage(c2_t<T> (&v)[P], c2_t<T>* lds, int lstride, int t,
                                          const c2_t<T>* __restrict__ tw, const StageTw<T, N, P, S>& ws) {
  using C = c2_t<T>;
  constexpr PlanDesc pd = make_plan(N, P);
  static_assert(pd.nst > 0, "unsupported FFT length");
  if constexpr (S < pd.nst) {
    constexpr int R = pd.r[S];
    constexpr int L = pd.L[S];
    constexpr int TPT = N / P;
    constexpr int NB = P / R;
    constexpr bool last = (S == pd.nst - 1);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int j = t + q * TPT;
      C a[R];
#pragma unroll
      for (int m = 0; m < R; ++m) a[m] = v[q + m * NB];
      if constexpr (S > 0) {
#pragma unroll
        for (int m = 1; m < R; ++m) {
          C wm = ws.w[q][m - 1];
          if (DIR > 0) wm.y = -wm.y;
          a[m] = cmul(a[m], wm);
        }
      }
      dft<R, DIR, T>(a);
      if constexpr (last) {
#pragma unroll
        for (int m = 0; m < R; ++m) v[q + m * NB] = a[m];
      } else {
        const int k = j % L;
        const int base = (j - k) * R + k;
        constexpr int PAD = 1 << PS;
        if constexpr (L % PAD == 0 || (L == 1 && PAD % R == 0)) {
          // lpos(base + m*L) = lpos(base) + m*L + (m*L >> PS): m*L is a
          // multiple of 2^PS, or base is a multiple of R and m < R divides 2^PS
          const int lb = lpos<PS>(base);
#pragma unroll
          for (int m = 0; m < R; ++m) lds[(lb + m * L + ((m * L) >> PS)) * lstride] = a[m];
        } else {
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lpos<PS>(base + m * L) * lstride] = a[m];
        }
      }
    }
    if constexpr (!last) {
      StageTw<T, N, P, S + 1> wn;
      load_tw<T, N, P, S + 1>(wn, t, tw);
      __syncthreads();
      if constexpr (TPT % (1 << PS) == 0) {
        const int lt = lpos<PS>(t);
#pragma unroll
        for (int i = 0; i < P; ++i) v[i] = lds[(lt + i * (TPT + (TPT >> PS))) * lstride];
      } else {
#pragma unroll
        for (int i = 0; i < P; ++i) v[i] = lds[lpos<PS>(t + i * TPT) * lstride];
      }
      __syncthreads();
      fft_stage<T, N, P, DIR, S + 1, PS>(v, lds, lstride, t, tw, wn);
    }
  }
}

template <typename T, int N, int P, int DIR, int PS = 3>
__device__ __forceinline__ void fft_stages(c2_t<T> (&v)[P], c2_t<T>* lds, int lstride, int t,
                                           const c2_t<T>* __restrict__ tw) {
  StageTw<T, N, P, 0> w0;
  fft_stage<T, N, P, DIR, 0, PS>(v, lds, lstride, t, tw, w0);
}

// Wavefront + workgroup sum of two doubles; result valid in thread 0.
// `red` must hold 2 * (blockDim/64) doubles of LDS.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0, sb = 0;
    for (int i = 0; i < nw; ++i) {
      sa += red[2 * i];
      sb += red[2 * i + 1];
    }
    a = sa;
    b = sb;
  }
}

// BaseCDI::update_n_best's ranking (NADIA/src/BaseCDI.c++:390-413), one
// thread: error e ranks against the n kept errors (ascending); on a place,
// the buffer of the worst entry becomes the new one (perm maps rank ->
// buffer) and *target names it for the snapshot copy, else *target = -1.  A
// NaN error ranks first, as in the reference's `>` test.
__device__ __forceinline__ void nbest_rank(double e, double* best_err, int* perm, int n, int* target) {
  int place = 0;
  while (place < n && e > best_err[place]) place++;
  if (place < n) {
    const int t = perm[n - 1];
    for (int i = n - 1; i > place; i--) {
      best_err[i] = best_err[i - 1];
      perm[i] = perm[i - 1];
    }
    best_err[place] = e;
    perm[place] = t;
    *target = t;
  } else {
    *target = -1;
  }
}

// The error of an iteration from n partial (diff, norm) pairs: the first 256
// threads of the workgroup sum them in a fixed order and the rest add
// zeros, so the value does not depend on which kernel (and workgroup size)
// reduces it.  Result in out[0]; `red` as for block_sum2.
__device__ __forceinline__ void finalize_err256(const double* p, int n, double* out, double* red) {
  double d = 0, m = 0;
  if (threadIdx.x < 256)
    for (int i = threadIdx.x; i < n; i += 256) {
      d += p[2 * i];
      m += p[2 * i + 1];
    }
  block_sum2(d, m, red);
  if (threadIdx.x == 0) out[0] = d / m;
}

__device__ __forceinline__ double block_max(double a, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(a, off, 64);
    a = o > a ? o : a;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) red[w] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int i = 1; i < nw; ++i) m = red[i] > m ? red[i] : m;
    a = m;
  }
  return a;
}

}  // namespace nadia
