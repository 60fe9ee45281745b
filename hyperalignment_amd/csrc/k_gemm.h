// k_gemm.h -- workgroup-level bf16x3 MFMA GEMM engine for gfx950.
//
// One 256-thread workgroup computes a full kp x kp product C = A B (or only
// the tiles I <= J of a symmetric product) tile by tile, with
//   * TILE x TILE output tiles, each wave owning a (TILE/2)^2 quarter made of
//     32x32 blocks of v_mfma_f32_32x32x16_bf16 (FP32 accumulation);
//   * bf16x3 products: C += Bhi*Ahi + Blo*Ahi + Bhi*Alo (the Blo*Alo term,
//     2^-16 relative, is dropped), so FP32-level accuracy at bf16 MFMA rate;
//   * K chunks of 32 staged in double-buffered LDS: the next chunk's global
//     loads are in flight while the current chunk's MFMAs run; k-contiguous
//     operands are stored as 16-byte slots of 8 k-values per row m, and an
//     m-contiguous A operand (OP_SPLIT_M) as k rows of m-contiguous bf16 whose
//     MFMA fragments come from transposing LDS reads (ds_read_b64_tr_b16), so
//     every staging write is a 128-bit store free of bank conflicts;
//   * 32x32 blocks that lie entirely outside the valid rows/columns skip their
//     MFMAs (edge tiles of kp that is not a multiple of TILE).
//
// Operand kinds (element (m, k) of the operand):
//   OP_F32_GATHER  f32[col(m) * ld + k], col(m) = cols[m] (cols == nullptr:
//                  col(m) = m): FP32 data, k contiguous
//   OP_SPLIT_K     hi/lo[m * ld + k]: pre-split planes, k contiguous
//   OP_SPLIT_GATHER hi/lo[cols[m] * ld + k]: pre-split planes, k contiguous,
//                  rows gathered (the row offsets are read once per tile)
//   OP_SPLIT_M     hi/lo[m + k * ld]: pre-split planes, m contiguous
//   OP_F32_M       f32[m + k * ld]: FP32 data, m contiguous
// Rows m >= mlimit read as zero; every k < K is in bounds.
//
// Two precisions share the staging buffers (PREC template argument):
//   PREC_BF16X3  FP32 operands are split into bf16 hi/lo while staging and
//                every product runs as the three bf16 MFMAs above;
//   PREC_F32     operands stay FP32 (only the FP32 kinds) and every product
//                runs on v_mfma_f32_32x32x2_f32: FP32 products at a sixth of
//                the bf16x3 rate, for paths whose conditioning needs them.
//
// MFMA operand order is swapped (the B fragment feeds the instruction's A
// slot), so lane % 32 indexes the output row m and output columns are
// contiguous per register: column-major stores coalesce.
#ifndef K_GEMM_H
#define K_GEMM_H

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) bf16x4_t lds_bf16x4_t;
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

#define NT 256  // threads per workgroup (4 waves)
#define KC 32   // K chunk staged in LDS

enum { OP_F32_GATHER = 0, OP_SPLIT_K = 1, OP_SPLIT_M = 2, OP_F32_M = 3, OP_SPLIT_GATHER = 4 };
enum { PREC_BF16X3 = 0, PREC_F32 = 1 };

// ---- bf16 helpers (round to nearest even) ----
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

struct SplitM {
  uint16_t *hi, *lo;
};
// a split matrix occupies the bytes of a kp x kp FP32 buffer: hi plane then lo plane
__device__ __forceinline__ SplitM as_split(float *buf, int kp) {
  SplitM s;
  s.hi = (uint16_t *)buf;
  s.lo = s.hi + (int64_t)kp * kp;
  return s;
}
__device__ __forceinline__ void put_split(const SplitM &s, int64_t i, float v) {
  uint16_t h = f2bf(v);
  s.hi[i] = h;
  s.lo[i] = f2bf(v - bf2f(h));
}

// Device-scope loads (cache policy sc1) of memory that other workgroups of the
// same launch write between barriers (the members of a team): they skip the
// CU's vector L1 and read the L2, so a barrier needs no L1 invalidate. They are
// buffer loads: base must be the same in every lane (the buffer descriptor
// lives in scalar registers) and byte offsets stay below 2^31. COH = false
// gives the plain load.
constexpr int CPOL_SC1 = 16;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t dev_rsrc(const void *base) {
  // word 3 = 0x00020000: the raw (unformatted) buffer word of gfx9-family descriptors
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, 0x7fffffff, 0x00020000);
}
template <bool COH>
__device__ __forceinline__ uint4 ld_b128(const void *base, int64_t ofs) {
  if constexpr (COH)
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(dev_rsrc(base), (int)ofs, 0, CPOL_SC1));
  else
    return *reinterpret_cast<const uint4 *>(static_cast<const char *>(base) + ofs);
}
template <bool COH>
__device__ __forceinline__ uint2 ld_b64(const void *base, int64_t ofs) {
  if constexpr (COH)
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(dev_rsrc(base), (int)ofs, 0, CPOL_SC1));
  else
    return *reinterpret_cast<const uint2 *>(static_cast<const char *>(base) + ofs);
}
template <bool COH>
__device__ __forceinline__ uint32_t ld_b32(const void *base, int64_t ofs) {
  if constexpr (COH)
    return __builtin_amdgcn_raw_buffer_load_b32(dev_rsrc(base), (int)ofs, 0, CPOL_SC1);
  else
    return *reinterpret_cast<const uint32_t *>(static_cast<const char *>(base) + ofs);
}
template <bool COH>
__device__ __forceinline__ uint16_t ld_b16(const void *base, int64_t ofs) {
  if constexpr (COH)
    return __builtin_amdgcn_raw_buffer_load_b16(dev_rsrc(base), (int)ofs, 0, CPOL_SC1);
  else
    return *reinterpret_cast<const uint16_t *>(static_cast<const char *>(base) + ofs);
}

// 16-byte load through a global-address-space pointer (a generic pointer
// would give a flat load, which also waits on the LDS counter)
__device__ __forceinline__ uint4 gld16(const uint16_t *p) {
  typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(1))) const u32x4_t gu32x4_t;
  return __builtin_bit_cast(uint4, *(gu32x4_t *)(p));
}

template <bool COH = false>
__device__ __forceinline__ float get_split(const SplitM &s, int64_t i) {
  return bf2f(ld_b16<COH>(s.hi, 2 * i)) + bf2f(ld_b16<COH>(s.lo, 2 * i));
}
// four consecutive elements i..i+3 (i a multiple of 4): one 8-byte access per plane
__device__ __forceinline__ void put_split4(const SplitM &s, int64_t i, float4 v) {
  const float z[4] = {v.x, v.y, v.z, v.w};
  uint16_t h[4], l[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    h[c] = f2bf(z[c]);
    l[c] = f2bf(z[c] - bf2f(h[c]));
  }
  *reinterpret_cast<uint2 *>(s.hi + i) =
      make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
  *reinterpret_cast<uint2 *>(s.lo + i) =
      make_uint2((uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16));
}
template <bool COH = false>
__device__ __forceinline__ float4 get_split4(const SplitM &s, int64_t i) {
  const uint2 h = ld_b64<COH>(s.hi, 2 * i), l = ld_b64<COH>(s.lo, 2 * i);
  return make_float4(bf2f((uint16_t)h.x) + bf2f((uint16_t)l.x), bf2f((uint16_t)(h.x >> 16)) + bf2f((uint16_t)(l.x >> 16)),
                     bf2f((uint16_t)h.y) + bf2f((uint16_t)l.y), bf2f((uint16_t)(h.y >> 16)) + bf2f((uint16_t)(l.y >> 16)));
}

struct Opnd {
  int32_t kind;
  const float *f32;
  const int32_t *cols;
  const uint16_t *hi, *lo;
  int32_t ld;
  int32_t mlimit;
};

__device__ __forceinline__ Opnd opnd_gather(const float *base, const int32_t *cols, int ld, int mlimit) {
  Opnd o;
  o.kind = OP_F32_GATHER; o.f32 = base; o.cols = cols; o.hi = nullptr; o.lo = nullptr; o.ld = ld; o.mlimit = mlimit;
  return o;
}
// pre-split planes (split_bf16) with rows gathered through cols
__device__ __forceinline__ Opnd opnd_gather_split(const uint16_t *hi, const uint16_t *lo, const int32_t *cols, int ld,
                                                  int mlimit) {
  Opnd o;
  o.kind = OP_SPLIT_GATHER; o.f32 = nullptr; o.cols = cols; o.hi = hi; o.lo = lo; o.ld = ld; o.mlimit = mlimit;
  return o;
}
__device__ __forceinline__ Opnd opnd_f32(const float *base, int kind, int ld, int mlimit) {
  Opnd o;
  o.kind = kind; o.f32 = base; o.cols = nullptr; o.hi = nullptr; o.lo = nullptr; o.ld = ld; o.mlimit = mlimit;
  return o;
}
__device__ __forceinline__ Opnd opnd_split(const SplitM &s, int kind, int ld, int mlimit) {
  Opnd o;
  o.kind = kind; o.f32 = nullptr; o.cols = nullptr; o.hi = s.hi; o.lo = s.lo; o.ld = ld; o.mlimit = mlimit;
  return o;
}

// LDS row strides. A k-contiguous row holds TILE slots (8 k-values of one m
// each, 16 bytes) padded to TILE + 2: 128-bit LDS accesses are served 8 lanes
// at a time, and the 8 lanes of a staging write (4 k-slots of 2 rows m) then
// fall in 8 distinct 16-byte bank groups. An m-contiguous row holds TILE 16-bit
// values padded by 16: a row stride of 32 bytes modulo 128, so the 4 rows of
// one transposing read (16 lanes, 4 rows x 32 bytes) fall in distinct banks.
template <int TILE>
struct Lds {
  static constexpr int SR = TILE + 2;
  static constexpr int TR = TILE + 16;
};
template <int TILE>
using SRow = uint4[Lds<TILE>::SR];

template <int TILE>
struct GemmSmem {
  union {
    struct {
      union {
        struct {
          uint4 ah[2][KC / 8][Lds<TILE>::SR], al[2][KC / 8][Lds<TILE>::SR];
        };
        struct {  // A operand of kind OP_SPLIT_M: k rows of m-contiguous bf16
          uint16_t th[2][KC][Lds<TILE>::TR], tl[2][KC][Lds<TILE>::TR];
        };
      };
      uint4 bh[2][KC / 8][Lds<TILE>::SR], bl[2][KC / 8][Lds<TILE>::SR];
    };
    float ctile[TILE * (TILE + 1)];  // finished tile, column-major, padded leading dim
  };
};

template <int TILE>
struct Stage {
  static constexpr int N = TILE * KC / (4 * NT);
  uint4 r[N];
};

// COH: device-scope loads, for operands that other workgroups of the launch
// wrote (not for gathered split rows; byte offsets below 2^31).
template <int TILE, int KIND, bool COH = false>
__device__ __forceinline__ void fetch(const Opnd &o, int m0, int kb, Stage<TILE> &st) {
  static_assert(!COH || KIND != OP_SPLIT_GATHER, "no device-scope loads for gathered split rows");
  constexpr int N = Stage<TILE>::N;
  const int tid = threadIdx.x;
  if constexpr (KIND == OP_SPLIT_GATHER) {
#pragma unroll
    for (int it = 0; it < N / 2; ++it) {
      const int e = tid + it * NT;
      const int m = e / (KC / 8), kg = e % (KC / 8);
      const int gm = m0 + m;
      uint4 h = make_uint4(0, 0, 0, 0), l = h;
      if (gm < o.mlimit) {
        const int64_t ofs = (int64_t)o.cols[gm] * o.ld + kb + kg * 8;
        h = gld16(o.hi + ofs);
        l = gld16(o.lo + ofs);
      }
      st.r[it] = h;
      st.r[N / 2 + it] = l;
    }
  } else if constexpr (KIND == OP_F32_GATHER) {
#pragma unroll
    for (int it = 0; it < N; ++it) {
      const int e = tid + it * NT;
      const int m = e / (KC / 4), kg = e % (KC / 4);
      const int gm = m0 + m;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (gm < o.mlimit) v = ld_b128<COH>(o.f32, 4 * ((int64_t)(o.cols ? o.cols[gm] : gm) * o.ld + kb + kg * 4));
      st.r[it] = v;
    }
  } else if constexpr (KIND == OP_F32_M) {
#pragma unroll
    for (int it = 0; it < N; ++it) {
      const int e = tid + it * NT;
      const int kk = e / (TILE / 4), mq = e % (TILE / 4);
      const int gm = m0 + mq * 4;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (gm < o.mlimit) v = ld_b128<COH>(o.f32, 4 * (gm + (int64_t)(kb + kk) * o.ld));  // mlimit % 4 == 0
      st.r[it] = v;
    }
  } else if constexpr (KIND == OP_SPLIT_K) {
#pragma unroll
    for (int it = 0; it < N / 2; ++it) {
      const int e = tid + it * NT;
      const int m = e / (KC / 8), kg = e % (KC / 8);
      const int gm = m0 + m;
      uint4 h = make_uint4(0, 0, 0, 0), l = h;
      if (gm < o.mlimit) {
        const int64_t ofs = (int64_t)gm * o.ld + kb + kg * 8;
        h = ld_b128<COH>(o.hi, 2 * ofs);
        l = ld_b128<COH>(o.lo, 2 * ofs);
      }
      st.r[it] = h;
      st.r[N / 2 + it] = l;
    }
  } else {
#pragma unroll
    for (int it = 0; it < N / 2; ++it) {
      const int e = tid + it * NT;
      const int kk = e / (TILE / 8), mg = e % (TILE / 8);
      const int gm = m0 + mg * 8;
      uint4 h = make_uint4(0, 0, 0, 0), l = h;
      if (gm < o.mlimit) {  // mlimit is a multiple of 8 here
        const int64_t ofs = gm + (int64_t)(kb + kk) * o.ld;
        h = ld_b128<COH>(o.hi, 2 * ofs);
        l = ld_b128<COH>(o.lo, 2 * ofs);
      }
      st.r[it] = h;
      st.r[N / 2 + it] = l;
    }
  }
}

// OP_SPLIT_GATHER rows of one tile: each thread's element offsets cols[m] * ld
// (-1: row m outside the operand), the same for every K chunk of the tile
template <int TILE>
__device__ __forceinline__ void gather_rows(const Opnd &o, int m0, int64_t *rows) {
#pragma unroll
  for (int it = 0; it < Stage<TILE>::N / 2; ++it) {
    const int gm = m0 + (threadIdx.x + it * NT) / (KC / 8);
    rows[it] = gm < o.mlimit ? (int64_t)o.cols[gm] * o.ld : -1;
  }
}
template <int TILE>
__device__ __forceinline__ void fetch_rows(const Opnd &o, const int64_t *rows, int kb, Stage<TILE> &st) {
  constexpr int N = Stage<TILE>::N;
#pragma unroll
  for (int it = 0; it < N / 2; ++it) {
    const int kg = (threadIdx.x + it * NT) % (KC / 8);
    uint4 h = make_uint4(0, 0, 0, 0), l = h;
    if (rows[it] >= 0) {
      const int64_t ofs = rows[it] + kb + kg * 8;
      h = gld16(o.hi + ofs);
      l = gld16(o.lo + ofs);
    }
    st.r[it] = h;
    st.r[N / 2 + it] = l;
  }
}
// one K chunk of an operand inside a tile (gathered rows from their offsets)
template <int TILE, int KIND, bool COH>
__device__ __forceinline__ void fetch_in_tile(const Opnd &o, int m0, const int64_t *rows, int kb, Stage<TILE> &st) {
  if constexpr (KIND == OP_SPLIT_GATHER)
    fetch_rows<TILE>(o, rows, kb, st);
  else
    fetch<TILE, KIND, COH>(o, m0, kb, st);
}

__device__ __forceinline__ uint32_t pack2(uint16_t a, uint16_t b) { return (uint32_t)a | ((uint32_t)b << 16); }

// FP32 staging view of one operand buffer: k quad q (k = 4q .. 4q+3) of row m
// lives in the hi array for q < KC/8 and in the lo array above.
template <int TILE>
__device__ __forceinline__ float4 *f32_quad(SRow<TILE> *sh, SRow<TILE> *sl, int q, int m) {
  return reinterpret_cast<float4 *>(q < KC / 8 ? &sh[q][m] : &sl[q - KC / 8][m]);
}

template <int TILE, int KIND, int PREC = PREC_BF16X3>
__device__ __forceinline__ void stash(const Stage<TILE> &st, SRow<TILE> *sh, SRow<TILE> *sl) {
  constexpr int N = Stage<TILE>::N;
  const int tid = threadIdx.x;
  if constexpr (PREC == PREC_F32) {
    static_assert(KIND == OP_F32_GATHER || KIND == OP_F32_M, "FP32 products take FP32 operands");
#pragma unroll
    for (int it = 0; it < N; ++it) {
      const int e = tid + it * NT;
      const uint4 u = st.r[it];
      if constexpr (KIND == OP_F32_GATHER) {
        const int m = e / (KC / 4), kg = e % (KC / 4);
        *f32_quad<TILE>(sh, sl, kg, m) = make_float4(__uint_as_float(u.x), __uint_as_float(u.y),
                                                     __uint_as_float(u.z), __uint_as_float(u.w));
      } else {
        const int kk = e / (TILE / 4), mq = e % (TILE / 4);
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int p = 0; p < 4; ++p)
          reinterpret_cast<float *>(f32_quad<TILE>(sh, sl, kk >> 2, mq * 4 + p))[kk & 3] = __uint_as_float(w[p]);
      }
    }
  } else if constexpr (KIND == OP_F32_GATHER) {
#pragma unroll
    for (int it = 0; it < N; ++it) {
      const int e = tid + it * NT;
      const int m = e / (KC / 4), kg = e % (KC / 4);
      const uint4 u = st.r[it];
      const float v0 = __uint_as_float(u.x), v1 = __uint_as_float(u.y);
      const float v2 = __uint_as_float(u.z), v3 = __uint_as_float(u.w);
      const uint16_t h0 = f2bf(v0), h1 = f2bf(v1), h2 = f2bf(v2), h3 = f2bf(v3);
      const uint16_t l0 = f2bf(v0 - bf2f(h0)), l1 = f2bf(v1 - bf2f(h1));
      const uint16_t l2 = f2bf(v2 - bf2f(h2)), l3 = f2bf(v3 - bf2f(h3));
      ((uint2 *)&sh[kg >> 1][m])[kg & 1] = make_uint2(pack2(h0, h1), pack2(h2, h3));
      ((uint2 *)&sl[kg >> 1][m])[kg & 1] = make_uint2(pack2(l0, l1), pack2(l2, l3));
    }
  } else if constexpr (KIND == OP_SPLIT_K || KIND == OP_SPLIT_GATHER) {
#pragma unroll
    for (int it = 0; it < N / 2; ++it) {
      const int e = tid + it * NT;
      const int m = e / (KC / 8), kg = e % (KC / 8);
      sh[kg][m] = st.r[it];
      sl[kg][m] = st.r[N / 2 + it];
    }
  } else {
    static_assert(KIND != OP_SPLIT_M && KIND != OP_F32_M,
                  "OP_SPLIT_M operands are staged by stash_m; OP_F32_M needs PREC_F32");
  }
}

// Stages an OP_SPLIT_M operand chunk (8 consecutive m of one k per 16 bytes)
// as k rows of m-contiguous bf16: 8 consecutive lanes write one 128-byte run.
template <int TILE>
__device__ __forceinline__ void stash_m(const Stage<TILE> &st, uint16_t (*th)[Lds<TILE>::TR],
                                        uint16_t (*tl)[Lds<TILE>::TR]) {
  constexpr int N = Stage<TILE>::N;
  const int tid = threadIdx.x;
#pragma unroll
  for (int it = 0; it < N / 2; ++it) {
    const int e = tid + it * NT;
    const int kk = e / (TILE / 8), mg = e % (TILE / 8);
    *reinterpret_cast<uint4 *>(&th[kk][mg * 8]) = st.r[it];
    *reinterpret_cast<uint4 *>(&tl[kk][mg * 8]) = st.r[N / 2 + it];
  }
}

// acc.v[i][j] holds C(m, n) for m = tm0 + wm0 + i*32 + lane%32 and
// n = tn0 + wn0 + j*32 + (r/4)*8 + (lane/32)*4 + r%4 (r = element index).
template <int TILE>
struct Acc {
  static constexpr int W = TILE / 64;
  f32x16_t v[W][W];
};

// MFMA fragment (8 consecutive k from k0, row mrow0 + lane % 16) from k rows of
// m-contiguous bf16 by two ds_read_b64_tr_b16. Within each 16-lane group that
// instruction hands lane i the 4 values M[4j + i/4][i % 4] (j = 0..3), where
// M[r] is the 64-bit piece addressed by lane r: lane r therefore addresses row
// k0 + r/4 at column mrow0 + 4 (r % 4), and lane i receives rows k0 .. k0+3 of
// column mrow0 + i.
template <int TILE>
__device__ __forceinline__ uint4 frag_m(const uint16_t (*t)[Lds<TILE>::TR], int k0, int mrow0, int lane) {
  const int r = lane & 15;
  uint16_t *p0 = const_cast<uint16_t *>(&t[k0 + (r >> 2)][mrow0 + 4 * (r & 3)]);
  uint16_t *p1 = const_cast<uint16_t *>(&t[k0 + 4 + (r >> 2)][mrow0 + 4 * (r & 3)]);
  const uint2 a = __builtin_bit_cast(uint2, __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4_t *)p0));
  const uint2 b = __builtin_bit_cast(uint2, __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4_t *)p1));
  return make_uint4(a.x, a.y, b.x, b.y);
}

template <int TILE, int KA = OP_SPLIT_K>
__device__ __forceinline__ void mma_chunk(const GemmSmem<TILE> &sm, int buf, Acc<TILE> &acc, const bool *bm,
                                          const bool *bn) {
  constexpr int W = TILE / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm0 = (wave >> 1) * (TILE / 2), wn0 = (wave & 1) * (TILE / 2);
#pragma unroll
  for (int ks = 0; ks < KC / 16; ++ks) {
    const int g = ks * 2 + (lane >> 5);
    bf16x8_t ahf[W], alf[W], bhf[W], blf[W];
#pragma unroll
    for (int i = 0; i < W; ++i) {
      uint4 h, l;
      if constexpr (KA == OP_SPLIT_M) {
        const int k0 = ks * 16 + 8 * (lane >> 5), mrow0 = wm0 + i * 32 + 16 * ((lane >> 4) & 1);
        h = frag_m<TILE>(sm.th[buf], k0, mrow0, lane);
        l = frag_m<TILE>(sm.tl[buf], k0, mrow0, lane);
      } else {
        h = sm.ah[buf][g][wm0 + i * 32 + (lane & 31)];
        l = sm.al[buf][g][wm0 + i * 32 + (lane & 31)];
      }
      ahf[i] = __builtin_bit_cast(bf16x8_t, h);
      alf[i] = __builtin_bit_cast(bf16x8_t, l);
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
      uint4 h = sm.bh[buf][g][wn0 + j * 32 + (lane & 31)];
      uint4 l = sm.bl[buf][g][wn0 + j * 32 + (lane & 31)];
      bhf[j] = __builtin_bit_cast(bf16x8_t, h);
      blf[j] = __builtin_bit_cast(bf16x8_t, l);
    }
#pragma unroll
    for (int i = 0; i < W; ++i)
#pragma unroll
      for (int j = 0; j < W; ++j) {
        if (!(bm[i] && bn[j])) continue;
        acc.v[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bhf[j], ahf[i], acc.v[i][j], 0, 0, 0);
        acc.v[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(blf[j], ahf[i], acc.v[i][j], 0, 0, 0);
        acc.v[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bhf[j], alf[i], acc.v[i][j], 0, 0, 0);
      }
  }
}

// FP32 products: v_mfma_f32_32x32x2_f32 takes k = 2 s + lane/32 of step s, so
// one float4 quad serves two steps (components lane/32 and 2 + lane/32).
template <int TILE>
__device__ __forceinline__ void mma_chunk_f32(GemmSmem<TILE> &sm, int buf, Acc<TILE> &acc, const bool *bm,
                                              const bool *bn) {
  constexpr int W = TILE / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm0 = (wave >> 1) * (TILE / 2), wn0 = (wave & 1) * (TILE / 2);
  const int h = lane >> 5;
#pragma unroll
  for (int q = 0; q < KC / 4; ++q) {
    float a0[W], a1[W], b0[W], b1[W];
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const float4 av = *f32_quad<TILE>(sm.ah[buf], sm.al[buf], q, wm0 + i * 32 + (lane & 31));
      a0[i] = h ? av.y : av.x;
      a1[i] = h ? av.w : av.z;
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const float4 bv = *f32_quad<TILE>(sm.bh[buf], sm.bl[buf], q, wn0 + j * 32 + (lane & 31));
      b0[j] = h ? bv.y : bv.x;
      b1[j] = h ? bv.w : bv.z;
    }
#pragma unroll
    for (int i = 0; i < W; ++i)
#pragma unroll
      for (int j = 0; j < W; ++j) {
        if (!(bm[i] && bn[j])) continue;
        acc.v[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(b0[j], a0[i], acc.v[i][j], 0, 0, 0);
        acc.v[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(b1[j], a1[i], acc.v[i][j], 0, 0, 0);
      }
  }
}

// acc += dadd * A(m, tn0 + n) for the columns n of the tile that the K chunk in
// buffer buf covers (kn = its first k minus tn0, a multiple of 32): over all
// chunks the tile becomes that of A (B + dadd I), from the A rows already in
// LDS. Only the waves whose 32-column block is kn do work.
template <int TILE, int KA, int PREC>
__device__ __forceinline__ void add_shift_chunk(GemmSmem<TILE> &sm, int buf, int kn, float dadd, Acc<TILE> &acc) {
  static_assert(KA == OP_SPLIT_K || (PREC == PREC_F32 && KA == OP_F32_GATHER), "k-contiguous A operand");
  constexpr int W = TILE / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm0 = (wave >> 1) * (TILE / 2), wn0 = (wave & 1) * (TILE / 2);
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (wn0 + j * 32 != kn) continue;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const int m = wm0 + i * 32 + (lane & 31);
      // element r of the block is column (r / 4) * 8 + (lane / 32) * 4 + r % 4
#pragma unroll
      for (int g = 0; g < KC / 8; ++g) {
        float v[4];
        if constexpr (PREC == PREC_F32) {
          const float4 q = *f32_quad<TILE>(sm.ah[buf], sm.al[buf], 2 * g + (lane >> 5), m);
          v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
          const uint2 h = reinterpret_cast<const uint2 *>(&sm.ah[buf][g][m])[lane >> 5];
          const uint2 l = reinterpret_cast<const uint2 *>(&sm.al[buf][g][m])[lane >> 5];
          v[0] = bf2f((uint16_t)h.x) + bf2f((uint16_t)l.x);
          v[1] = bf2f((uint16_t)(h.x >> 16)) + bf2f((uint16_t)(l.x >> 16));
          v[2] = bf2f((uint16_t)h.y) + bf2f((uint16_t)l.y);
          v[3] = bf2f((uint16_t)(h.y >> 16)) + bf2f((uint16_t)(l.y >> 16));
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) acc.v[i][j][4 * g + t] += dadd * v[t];
      }
    }
  }
}

// One output tile over K. sa, sb: the stage registers; with have_first the
// caller has already fetched chunk 0 of this tile into them (during the
// previous tile's epilogue), so its load latency is hidden. dadd != 0: the
// tile of A (B + dadd I) instead (square products, k-contiguous A).
template <int TILE, int KA, int KB, int PREC = PREC_BF16X3, bool COH = false>
__device__ __forceinline__ void tile_gemm(GemmSmem<TILE> &sm, const Opnd &A, const Opnd &B, int tm0, int tn0,
                                          int K, Acc<TILE> &acc, Stage<TILE> &sa, Stage<TILE> &sb,
                                          bool have_first, float dadd = 0.f) {
  constexpr int W = TILE / 64;
  const int wave = threadIdx.x >> 6;
  const int wm0 = (wave >> 1) * (TILE / 2), wn0 = (wave & 1) * (TILE / 2);
  bool bm[W], bn[W];
#pragma unroll
  for (int i = 0; i < W; ++i) {
    bm[i] = tm0 + wm0 + i * 32 < A.mlimit;
    bn[i] = tn0 + wn0 + i * 32 < B.mlimit;
  }
#pragma unroll
  for (int i = 0; i < W; ++i)
#pragma unroll
    for (int j = 0; j < W; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc.v[i][j][r] = 0.f;
  static_assert(KB != OP_SPLIT_M, "only the A operand may be read with m contiguous");
  const int nc = K / KC;
  int64_t ra[Stage<TILE>::N / 2], rb[Stage<TILE>::N / 2];
  if constexpr (KA == OP_SPLIT_GATHER) gather_rows<TILE>(A, tm0, ra);
  if constexpr (KB == OP_SPLIT_GATHER) gather_rows<TILE>(B, tn0, rb);
  if (!have_first) {
    fetch_in_tile<TILE, KA, COH>(A, tm0, ra, 0, sa);
    fetch_in_tile<TILE, KB, COH>(B, tn0, rb, 0, sb);
  }
  if constexpr (KA == OP_SPLIT_M)
    stash_m<TILE>(sa, sm.th[0], sm.tl[0]);
  else
    stash<TILE, KA, PREC>(sa, sm.ah[0], sm.al[0]);
  stash<TILE, KB, PREC>(sb, sm.bh[0], sm.bl[0]);
  __syncthreads();
  for (int c = 0; c < nc; ++c) {
    const int cur = c & 1;
    const bool more = c + 1 < nc;
    if (more) {
      fetch_in_tile<TILE, KA, COH>(A, tm0, ra, (c + 1) * KC, sa);
      fetch_in_tile<TILE, KB, COH>(B, tn0, rb, (c + 1) * KC, sb);
    }
    if constexpr (PREC == PREC_F32)
      mma_chunk_f32<TILE>(sm, cur, acc, bm, bn);
    else
      mma_chunk<TILE, KA>(sm, cur, acc, bm, bn);
    if constexpr (KA == OP_SPLIT_K || (PREC == PREC_F32 && KA == OP_F32_GATHER)) {
      const int kn = c * KC - tn0;
      if (dadd != 0.f && kn >= 0 && kn < TILE) add_shift_chunk<TILE, KA, PREC>(sm, cur, kn, dadd, acc);
    }
    if (more) {
      if constexpr (KA == OP_SPLIT_M)
        stash_m<TILE>(sa, sm.th[cur ^ 1], sm.tl[cur ^ 1]);
      else
        stash<TILE, KA, PREC>(sa, sm.ah[cur ^ 1], sm.al[cur ^ 1]);
      stash<TILE, KB, PREC>(sb, sm.bh[cur ^ 1], sm.bl[cur ^ 1]);
    }
    __syncthreads();
  }
}

// Writes the accumulators into sm.ctile: element (ml, nl) of the tile at ctile[nl * (TILE + 1) + ml].
template <int TILE>
__device__ __forceinline__ void acc_to_lds(const Acc<TILE> &acc, float *ctile) {
  constexpr int W = TILE / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm0 = (wave >> 1) * (TILE / 2), wn0 = (wave & 1) * (TILE / 2);
#pragma unroll
  for (int i = 0; i < W; ++i)
#pragma unroll
    for (int j = 0; j < W; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ml = wm0 + i * 32 + (lane & 31);
        const int nl = wn0 + j * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        ctile[nl * (TILE + 1) + ml] = acc.v[i][j][r];
      }
}

// Full product over M x N (rows m < M from A, columns n < N from B; M and N are
// multiples of 32). Each finished tile is staged in LDS and handed to
// epi(m, n, v) four rows at a time: v holds C(m..m+3, n), m a multiple of 4.
// sym (M == N, A == B or a symmetric result): only tiles I <= J are computed
// and every off-diagonal tile is also delivered transposed, so epi still sees
// every (m, n) of the full product exactly once.
//
// Tile subsets (nw workgroups sharing one product, this one number w): tile
// (I, J) goes to workgroup (I + J) mod nw, each in row-major order. Every
// workgroup then has about one tile per row and per column, so the cheap edge
// tiles (kp not a multiple of TILE) spread evenly, and at any time the
// workgroups work on tiles of the same row and read its row panel together.
template <int TILE, int KA, int KB, int PREC = PREC_BF16X3, bool COH = false, class Epi>
__device__ void wg_gemm_mn(GemmSmem<TILE> &sm, const Opnd &A, const Opnd &B, int M, int N, int K, bool sym,
                           Epi epi, int nw = 1, int w = 0, float dadd = 0.f) {
  const int ntm = (M + TILE - 1) / TILE, ntn = (N + TILE - 1) / TILE;
  constexpr int LD = TILE + 1;
  // this workgroup's tiles in row-major order (only I <= J when sym)
  auto next_tile = [&](int &I, int &J) {
    for (;;) {
      if (++J >= ntn) {
        if (++I >= ntm) return false;
        J = sym ? I : 0;
      }
      if ((I + J) % nw == w) return true;
    }
  };
  Stage<TILE> sa, sb;
  bool have_first = false;
  int I = 0, J = -1;
  for (bool more = next_tile(I, J); more;) {
    {
      Acc<TILE> acc;
      tile_gemm<TILE, KA, KB, PREC, COH>(sm, A, B, I * TILE, J * TILE, K, acc, sa, sb, have_first, dadd);
      acc_to_lds<TILE>(acc, sm.ctile);
    }
    // the next tile's first K chunk is loaded while this tile is delivered
    int In = I, Jn = J;
    more = next_tile(In, Jn);
    if (more) {
      fetch<TILE, KA, COH>(A, In * TILE, 0, sa);
      fetch<TILE, KB, COH>(B, Jn * TILE, 0, sb);
    }
    have_first = more;
    __syncthreads();
    const int mmax = min(TILE, M - I * TILE), nmax = min(TILE, N - J * TILE);
    // four consecutive rows per call; 8 consecutive lanes take 8 row groups of
    // one column and the next lanes the next columns, so with LD = 1 mod 32 the
    // 32 lanes of an LDS read hit 32 distinct banks and each column's stores
    // are 64-byte runs per plane
    constexpr int TQ = TILE / 4;
    for (int e = threadIdx.x; e < TQ * TILE; e += NT) {
      const int rg = (e & 7) + 8 * (e / (8 * TILE)), nl = (e >> 3) % TILE;
      const int ml = 4 * rg;
      if (ml < mmax && nl < nmax) {
        const float *c = &sm.ctile[nl * LD + ml];
        epi(I * TILE + ml, J * TILE + nl, make_float4(c[0], c[1], c[2], c[3]));
      }
    }
    if (sym && I != J) {
      // transposed delivery: output (J*TILE + a, I*TILE + b) = tile element (b, a)
      for (int e = threadIdx.x; e < TQ * TILE; e += NT) {
        const int ag = (e & 7) + 8 * (e / (8 * TILE)), bl = (e >> 3) % TILE;
        const int al = 4 * ag;
        if (al < nmax && bl < mmax) {
          const float *c = &sm.ctile[al * LD + bl];
          epi(J * TILE + al, I * TILE + bl, make_float4(c[0], c[LD], c[2 * LD], c[3 * LD]));
        }
      }
    }
    __syncthreads();
    I = In;
    J = Jn;
  }
}

// Square kp x kp product (see wg_gemm_mn); dadd != 0: A (B + dadd I).
template <int TILE, int KA, int KB, int PREC = PREC_BF16X3, bool COH = false, class Epi>
__device__ __forceinline__ void wg_gemm(GemmSmem<TILE> &sm, const Opnd &A, const Opnd &B, int kp, int K, bool sym,
                                        Epi epi, int nw = 1, int w = 0, float dadd = 0.f) {
  wg_gemm_mn<TILE, KA, KB, PREC, COH>(sm, A, B, kp, kp, K, sym, epi, nw, w, dadd);
}

#endif  // K_GEMM_H
