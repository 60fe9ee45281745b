// k_index.hip -- integer and FP64 bookkeeping of the searchlight pass on
// gfx950: searchlight ownership of flat positions, the vertex->searchlight
// inverse index, distance weights (bit-identical to the reference FP64 loop),
// the CSC pattern, and the deterministic owner-computes scatter-add.
//
// Built with -ffp-contract=off: every FP64 expression here must round exactly
// like the reference's separate multiply and add.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hak.h"

#define LAUNCH_RET() return (int)hipGetLastError()

// ---------------------------------------------------------------------------
// fills
// ---------------------------------------------------------------------------

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_f64(double *p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static inline int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

extern "C" int hak_fill_i32(int32_t *p, int64_t n, int32_t v, void *stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, p, n, v);
  LAUNCH_RET();
}
extern "C" int hak_fill_f64(double *p, int64_t n, double v, void *stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, p, n, v);
  LAUNCH_RET();
}

// ---------------------------------------------------------------------------
// searchlight ownership and inverse index
// ---------------------------------------------------------------------------

__global__ void k_sl_of(const int32_t *__restrict__ sl_off, int32_t count, int32_t *__restrict__ sl_of) {
  for (int32_t s = blockIdx.x; s < count; s += gridDim.x) {
    int32_t b = sl_off[s], e = sl_off[s + 1];
    for (int32_t k = b + threadIdx.x; k < e; k += blockDim.x) sl_of[k] = s;
  }
}
extern "C" int hak_sl_of(const int32_t *sl_off, int32_t count, int32_t *sl_of, void *stream) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_sl_of, dim3(count < 65536 ? count : 65536), dim3(256), 0,
                     (hipStream_t)stream, sl_off, count, sl_of);
  LAUNCH_RET();
}

__global__ void k_count_occ(const int32_t *__restrict__ idx, int64_t total, int32_t *cnt) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[idx[k]], 1);
}
extern "C" int hak_count_occ(const int32_t *sl_idx, int64_t total, int32_t *cnt, void *stream) {
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_count_occ, dim3(grid_for(total, 256)), dim3(256), 0, (hipStream_t)stream,
                     sl_idx, total, cnt);
  LAUNCH_RET();
}

// Single-workgroup exclusive scan: 1024 threads, 4 items each per chunk, with
// a running carry. n is at most a few million here (vertices, searchlights).
#define SCAN_T 1024
#define SCAN_I 4
__global__ __launch_bounds__(SCAN_T) void k_scan_i32_i64(const int32_t *__restrict__ in, int64_t n,
                                                         int64_t *__restrict__ out) {
  __shared__ int64_t part[SCAN_T];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int64_t chunk = (int64_t)SCAN_T * SCAN_I;
  for (int64_t base = 0; base < n; base += chunk) {
    int64_t v[SCAN_I];
    int64_t local = 0;
    for (int i = 0; i < SCAN_I; ++i) {
      int64_t j = base + (int64_t)threadIdx.x * SCAN_I + i;
      v[i] = (j < n) ? (int64_t)in[j] : 0;
      local += v[i];
    }
    part[threadIdx.x] = local;
    __syncthreads();
    // Hillis-Steele inclusive scan over the per-thread sums
    for (int off = 1; off < SCAN_T; off <<= 1) {
      int64_t t = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    int64_t run = carry + part[threadIdx.x] - local;
    for (int i = 0; i < SCAN_I; ++i) {
      int64_t j = base + (int64_t)threadIdx.x * SCAN_I + i;
      if (j < n) out[j] = run;
      run += v[i];
    }
    __syncthreads();
    if (threadIdx.x == SCAN_T - 1) carry += part[SCAN_T - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}
extern "C" int hak_scan_i32_i64(const int32_t *in, int64_t n, int64_t *out, void *stream) {
  hipLaunchKernelGGL(k_scan_i32_i64, dim3(1), dim3(SCAN_T), 0, (hipStream_t)stream, in, n, out);
  LAUNCH_RET();
}

__global__ void k_i64_to_i32(const int64_t *__restrict__ in, int64_t n, int32_t *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (int32_t)in[i];
}
extern "C" int hak_i64_to_i32(const int64_t *in, int64_t n, int32_t *out, void *stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_i64_to_i32, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, in, n, out);
  LAUNCH_RET();
}

__global__ void k_fill_occ(const int32_t *__restrict__ idx, int64_t total, const int64_t *__restrict__ ptr,
                           int32_t *cursor, int32_t *__restrict__ occ) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = idx[k];
    int32_t p = atomicAdd(&cursor[v], 1);
    occ[ptr[v] + p] = (int32_t)k;
  }
}
extern "C" int hak_fill_occ(const int32_t *sl_idx, int64_t total, const int64_t *occ_ptr,
                            int32_t *cursor, int32_t *occ, void *stream) {
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_fill_occ, dim3(grid_for(total, 256)), dim3(256), 0, (hipStream_t)stream,
                     sl_idx, total, occ_ptr, cursor, occ);
  LAUNCH_RET();
}

// Per-vertex bitonic sort of the occurrence list in LDS (ascending flat
// position = ascending searchlight order). Lists longer than OCC_MAX set *err.
#define OCC_MAX 8192
__global__ __launch_bounds__(256) void k_sort_occ(const int64_t *__restrict__ ptr, int32_t nv,
                                                  const int32_t *__restrict__ sl_of,
                                                  int32_t *__restrict__ occ, int32_t *err) {
  __shared__ int32_t buf[OCC_MAX];
  for (int32_t v = blockIdx.x; v < nv; v += gridDim.x) {
    int64_t b = ptr[v];
    int32_t L = (int32_t)(ptr[v + 1] - b);
    if (L <= 1) continue;
    if (L > OCC_MAX) {
      if (threadIdx.x == 0) atomicExch(err, 1);
      continue;
    }
    int32_t P = 1;
    while (P < L) P <<= 1;
    for (int32_t i = threadIdx.x; i < P; i += blockDim.x) buf[i] = (i < L) ? occ[b + i] : 0x7fffffff;
    __syncthreads();
    for (int32_t size = 2; size <= P; size <<= 1) {
      for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (int32_t i = threadIdx.x; i < P; i += blockDim.x) {
          int32_t j = i ^ stride;
          if (j > i) {
            bool up = ((i & size) == 0);
            int32_t x = buf[i], y = buf[j];
            if ((x > y) == up) { buf[i] = y; buf[j] = x; }
          }
        }
        __syncthreads();
      }
    }
    for (int32_t i = threadIdx.x; i < L; i += blockDim.x) {
      occ[b + i] = buf[i];
      // the same vertex twice in one searchlight: consecutive entries share a searchlight
      if (i + 1 < L && sl_of[buf[i]] == sl_of[buf[i + 1]]) atomicExch(err, 2);
    }
    __syncthreads();
  }
}
extern "C" int hak_sort_occ(const int64_t *occ_ptr, int32_t nv, const int32_t *sl_of, int32_t *occ,
                            int32_t *err, void *stream) {
  if (nv <= 0) return 0;
  hipLaunchKernelGGL(k_sort_occ, dim3(nv < 65536 ? nv : 65536), dim3(256), 0, (hipStream_t)stream,
                     occ_ptr, nv, sl_of, occ, err);
  LAUNCH_RET();
}

// ---------------------------------------------------------------------------
// weights (ha_searchlight.c:970-1001): per-vertex sum in flat order, then divide
// ---------------------------------------------------------------------------

__device__ __forceinline__ double raw_weight(const double *dists, double radius, int64_t k) {
  return dists ? (radius - dists[k]) / radius : 1.0;
}

__global__ void k_weight_sum(const double *__restrict__ dists, double radius,
                             const int64_t *__restrict__ ptr, const int32_t *__restrict__ occ,
                             int32_t nv, double *__restrict__ wsum) {
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t p = ptr[v]; p < ptr[v + 1]; ++p) acc += raw_weight(dists, radius, occ[p]);
    wsum[v] = acc;
  }
}
__global__ void k_weight_div(const int32_t *__restrict__ idx, const double *__restrict__ dists,
                             double radius, int64_t total, const double *__restrict__ wsum,
                             double *__restrict__ w) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    double ws = wsum[idx[k]];
    w[k] = dists ? raw_weight(dists, radius, k) / ws : 1.0 / ws;
  }
}
extern "C" int hak_weights(const int32_t *sl_idx, const double *dists, double radius, int64_t total,
                           const int64_t *occ_ptr, const int32_t *occ, int32_t nv,
                           double *wsum, double *w, void *stream) {
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_weight_sum, dim3(grid_for(nv, 256)), dim3(256), 0, (hipStream_t)stream,
                     dists, radius, occ_ptr, occ, nv, wsum);
  hipLaunchKernelGGL(k_weight_div, dim3(grid_for(total, 256)), dim3(256), 0, (hipStream_t)stream,
                     sl_idx, dists, radius, total, wsum, w);
  LAUNCH_RET();
}

// ---------------------------------------------------------------------------
// CSC pattern
// ---------------------------------------------------------------------------

__device__ __forceinline__ int32_t block_reduce_min(int32_t v, int32_t *red) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6;
  int32_t r = red[0];
  for (int i = 1; i < nw; ++i) r = min(r, red[i]);
  __syncthreads();
  return r;
}
__device__ __forceinline__ int32_t block_reduce_max(int32_t v, int32_t *red) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6;
  int32_t r = red[0];
  for (int i = 1; i < nw; ++i) r = max(r, red[i]);
  __syncthreads();
  return r;
}
__device__ __forceinline__ int64_t block_reduce_sum64(int64_t v, int64_t *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6;
  int64_t r = 0;
  for (int i = 0; i < nw; ++i) r += red[i];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_sl_minmax(const int32_t *__restrict__ idx, const int32_t *__restrict__ off,
                                                   int32_t count, int32_t *slmin, int32_t *slmax) {
  __shared__ int32_t red[4];
  for (int32_t s = blockIdx.x; s < count; s += gridDim.x) {
    int32_t lo = 0x7fffffff, hi = -1;
    for (int32_t k = off[s] + threadIdx.x; k < off[s + 1]; k += blockDim.x) {
      lo = min(lo, idx[k]);
      hi = max(hi, idx[k]);
    }
    lo = block_reduce_min(lo, red);
    hi = block_reduce_max(hi, red);
    if (threadIdx.x == 0) { slmin[s] = lo; slmax[s] = hi; }
  }
}
extern "C" int hak_sl_minmax(const int32_t *sl_idx, const int32_t *sl_off, int32_t count,
                             int32_t *slmin, int32_t *slmax, void *stream) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_sl_minmax, dim3(count < 65536 ? count : 65536), dim3(256), 0,
                     (hipStream_t)stream, sl_idx, sl_off, count, slmin, slmax);
  LAUNCH_RET();
}

#define PAT_WORDS 8192   // 32 KB bitmap window = 262,144 row ids per window
#define PAT_T 256

// Marks into the LDS bitmap every row of every searchlight that contains
// column c, restricted to the window [w0, w0 + 32*PAT_WORDS).
__device__ void pattern_mark(uint32_t *bits, int32_t w0, const int32_t *__restrict__ idx,
                             const int32_t *__restrict__ off, const int32_t *__restrict__ sl_of,
                             int64_t pb, int64_t pe, const int32_t *__restrict__ occ) {
  for (int32_t i = threadIdx.x; i < PAT_WORDS; i += PAT_T) bits[i] = 0u;
  __syncthreads();
  const int32_t span = 32 * PAT_WORDS;
  for (int64_t p = pb; p < pe; ++p) {
    int32_t s = sl_of[occ[p]];
    int32_t b = off[s], e = off[s + 1];
    for (int32_t k = b + threadIdx.x; k < e; k += PAT_T) {
      int32_t r = idx[k] - w0;
      if (r >= 0 && r < span) atomicOr(&bits[r >> 5], 1u << (r & 31));
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(PAT_T) void k_pattern_count(
    const int32_t *__restrict__ idx, const int32_t *__restrict__ off, const int32_t *__restrict__ sl_of,
    const int64_t *__restrict__ ptr, const int32_t *__restrict__ occ, const int32_t *__restrict__ slmin,
    const int32_t *__restrict__ slmax, int32_t nv, int32_t *__restrict__ colcnt) {
  __shared__ uint32_t bits[PAT_WORDS];
  __shared__ int32_t red[4];
  __shared__ int64_t red64[4];
  for (int32_t c = blockIdx.x; c < nv; c += gridDim.x) {
    int64_t pb = ptr[c], pe = ptr[c + 1];
    int32_t lo = 0x7fffffff, hi = -1;
    for (int64_t p = pb + threadIdx.x; p < pe; p += PAT_T) {
      int32_t s = sl_of[occ[p]];
      lo = min(lo, slmin[s]);
      hi = max(hi, slmax[s]);
    }
    lo = block_reduce_min(lo, red);
    hi = block_reduce_max(hi, red);
    int64_t total = 0;
    for (int32_t w0 = lo; hi >= 0 && w0 <= hi; w0 += 32 * PAT_WORDS) {
      pattern_mark(bits, w0, idx, off, sl_of, pb, pe, occ);
      int64_t cnt = 0;
      for (int32_t i = threadIdx.x; i < PAT_WORDS; i += PAT_T) cnt += __popc(bits[i]);
      total += block_reduce_sum64(cnt, red64);
    }
    if (threadIdx.x == 0) colcnt[c] = (int32_t)total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(PAT_T) void k_pattern_fill(
    const int32_t *__restrict__ idx, const int32_t *__restrict__ off, const int32_t *__restrict__ sl_of,
    const int64_t *__restrict__ ptr, const int32_t *__restrict__ occ, const int32_t *__restrict__ slmin,
    const int32_t *__restrict__ slmax, int32_t nv, const int64_t *__restrict__ indptr,
    int32_t *__restrict__ indices) {
  __shared__ uint32_t bits[PAT_WORDS];
  __shared__ int32_t red[4];
  __shared__ int32_t scan[PAT_T];
  const int WPT = PAT_WORDS / PAT_T;  // words per thread (contiguous range)
  for (int32_t c = blockIdx.x; c < nv; c += gridDim.x) {
    int64_t pb = ptr[c], pe = ptr[c + 1];
    int32_t lo = 0x7fffffff, hi = -1;
    for (int64_t p = pb + threadIdx.x; p < pe; p += PAT_T) {
      int32_t s = sl_of[occ[p]];
      lo = min(lo, slmin[s]);
      hi = max(hi, slmax[s]);
    }
    lo = block_reduce_min(lo, red);
    hi = block_reduce_max(hi, red);
    int64_t outpos = indptr[c];
    for (int32_t w0 = lo; hi >= 0 && w0 <= hi; w0 += 32 * PAT_WORDS) {
      pattern_mark(bits, w0, idx, off, sl_of, pb, pe, occ);
      int32_t cnt = 0;
      for (int i = 0; i < WPT; ++i) cnt += __popc(bits[threadIdx.x * WPT + i]);
      scan[threadIdx.x] = cnt;
      __syncthreads();
      for (int o = 1; o < PAT_T; o <<= 1) {
        int32_t t = (threadIdx.x >= (unsigned)o) ? scan[threadIdx.x - o] : 0;
        __syncthreads();
        scan[threadIdx.x] += t;
        __syncthreads();
      }
      int64_t pos = outpos + scan[threadIdx.x] - cnt;
      for (int i = 0; i < WPT; ++i) {
        int32_t wi = threadIdx.x * WPT + i;
        uint32_t m = bits[wi];
        while (m) {
          int b = __ffs(m) - 1;
          m &= m - 1;
          indices[pos++] = w0 + wi * 32 + b;
        }
      }
      outpos += scan[PAT_T - 1];
      __syncthreads();
    }
  }
}

extern "C" int hak_pattern_count(const int32_t *sl_idx, const int32_t *sl_off, const int32_t *sl_of,
                                 const int64_t *occ_ptr, const int32_t *occ, const int32_t *slmin,
                                 const int32_t *slmax, int32_t nv, int32_t *colcnt, void *stream) {
  if (nv <= 0) return 0;
  hipLaunchKernelGGL(k_pattern_count, dim3(nv < 65536 ? nv : 65536), dim3(PAT_T), 0, (hipStream_t)stream,
                     sl_idx, sl_off, sl_of, occ_ptr, occ, slmin, slmax, nv, colcnt);
  LAUNCH_RET();
}
extern "C" int hak_pattern_fill(const int32_t *sl_idx, const int32_t *sl_off, const int32_t *sl_of,
                                const int64_t *occ_ptr, const int32_t *occ, const int32_t *slmin,
                                const int32_t *slmax, int32_t nv, const int64_t *indptr,
                                int32_t *indices, void *stream) {
  if (nv <= 0) return 0;
  hipLaunchKernelGGL(k_pattern_fill, dim3(nv < 65536 ? nv : 65536), dim3(PAT_T), 0, (hipStream_t)stream,
                     sl_idx, sl_off, sl_of, occ_ptr, occ, slmin, slmax, nv, indptr, indices);
  LAUNCH_RET();
}

// ---------------------------------------------------------------------------
// deterministic scatter (owner computes): one workgroup per CSC column
// ---------------------------------------------------------------------------

#define SC_T 256
#define SC_LDS 3072   // most column entries held in LDS (FP64 accumulators)
#define SC_HROW (1 << 19)  // rows below this fit the hash packing ((row + 1) << 12 | position)

// LDS hash of a column's rows: 2^hbits words of ((row + 1) << 12) | position,
// 0 = empty (positions < 4096 > SC_LDS), linear probing
__device__ __forceinline__ uint32_t sc_hash(int32_t r, int hbits) {
  return ((uint32_t)r * 2654435761u) >> (32 - hbits);
}

// position of row r in the column, from the LDS hash (-1: r is not in the column)
__device__ __forceinline__ int32_t hash_row(const uint32_t *ht, int hbits, int32_t r) {
  const uint32_t key = (uint32_t)(r + 1), mask = (1u << hbits) - 1u;
  uint32_t h = sc_hash(r, hbits);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const uint32_t e = ht[h];
    if (e == 0u) return -1;
    if ((e >> 12) == key) return (int32_t)(e & 0xFFFu);
    h = (h + 1) & mask;
  }
  return -1;
}

__device__ __forceinline__ int32_t find_row(const int32_t *rows, int32_t L, int32_t r) {
  int32_t lo = 0, hi = L;
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (rows[mid] < r) lo = mid + 1; else hi = mid;
  }
  return (lo < L && rows[lo] == r) ? lo : -1;
}

// Per column c (one workgroup), the contributions of every searchlight holding
// c, in increasing searchlight order with a barrier between searchlights, so
// each element of data is summed in one fixed order. A column of up to cap
// rows keeps its FP64 accumulators in LDS and finds a row's position through an
// LDS hash of 2^hbits words (about one probe; a row outside the pattern ends at
// an empty slot and is skipped); longer columns work in global memory with a
// binary search. cap is the plan's longest column (at most SC_LDS), so the
// dynamic LDS, and with it the workgroups per CU, fits the plan. The
// occurrences' metadata (searchlight, position, size, Q offset, scale) is
// staged in LDS a batch at a time by all threads at once, and the Q values,
// weights and vertex ids of SC_P searchlights are loaded together.
#define SC_OCC 64   // occurrences staged per batch
#define SC_V 2      // values per thread per searchlight held in registers (k <= SC_V * SC_T pipelined)
#define SC_P 8      // searchlights whose values are loaded together

// One thread's raw loads for one searchlight: Q values, weights, rows (-1: none).
// Only loads, no arithmetic: the arithmetic waits for its operands, and done
// here it would make each searchlight wait for its own loads.
struct ScRaw {
  float q[SC_V];
  double w[SC_V];
  int32_t r[SC_V];
};

__global__ __launch_bounds__(SC_T) void k_scatter(
    const int32_t *__restrict__ slx_idx, const int32_t *__restrict__ off, const int32_t *__restrict__ sl_of,
    const int64_t *__restrict__ ptr, const int32_t *__restrict__ occ, int32_t ncols,
    const int64_t *__restrict__ indptr, const int32_t *__restrict__ indices, double *__restrict__ data,
    const float *__restrict__ q, const int64_t *__restrict__ qoff, const int32_t *__restrict__ qld,
    const double *__restrict__ scale, const double *__restrict__ w, int32_t s_begin, int32_t s_end,
    int32_t cap, int32_t hbits) {
  // dynamic LDS: accumulators (cap), metadata (SC_OCC each), hash (2^hbits words)
  extern __shared__ __align__(16) unsigned char sc_lds[];
  double *acc_l = reinterpret_cast<double *>(sc_lds);
  int64_t *m_q = reinterpret_cast<int64_t *>(acc_l + cap);
  double *m_sc = reinterpret_cast<double *>(m_q + SC_OCC);
  int32_t *m_o = reinterpret_cast<int32_t *>(m_sc + SC_OCC);
  int32_t *m_ks = m_o + SC_OCC;
  uint32_t *ht = reinterpret_cast<uint32_t *>(m_ks + SC_OCC);
  const int32_t hsize = 1 << hbits;
  const int tid = threadIdx.x;
  for (int32_t c = blockIdx.x; c < ncols; c += gridDim.x) {
    // occurrences outside [s_begin, s_end) are skipped: occ lists are sorted
    // by flat position, i.e. by searchlight, so two binary searches find the
    // range (a linear walk costs two dependent loads per occurrence)
    auto first_at = [&](int64_t lo, int64_t hi, int32_t sv) {
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (sl_of[occ[mid]] < sv)
          lo = mid + 1;
        else
          hi = mid;
      }
      return lo;
    };
    const int64_t pb = first_at(ptr[c], ptr[c + 1], s_begin);
    const int64_t pe2 = first_at(pb, ptr[c + 1], s_end);
    if (pb == pe2) continue;
    const int64_t cb = indptr[c];
    const int32_t L = (int32_t)(indptr[c + 1] - cb);
    // rows sorted: the last is the largest
    const bool in_lds = L > 0 && L <= cap && indices[cb + L - 1] < SC_HROW;
    const int32_t *rows = indices + cb;
    double *acc = in_lds ? acc_l : data + cb;
    if (in_lds) {
      for (int32_t i = tid; i < hsize; i += SC_T) ht[i] = 0u;
      __syncthreads();
      for (int32_t i = tid; i < L; i += SC_T) {
        const int32_t r = indices[cb + i];
        const uint32_t v = ((uint32_t)(r + 1) << 12) | (uint32_t)i;
        uint32_t h = sc_hash(r, hbits);
        while (atomicCAS(&ht[h], 0u, v) != 0u) h = (h + 1) & (uint32_t)(hsize - 1);
        acc_l[i] = data[cb + i];
      }
    }
    auto pos_of = [&](int32_t r) { return in_lds ? hash_row(ht, hbits, r) : find_row(rows, L, r); };
    for (int64_t b0 = pb; b0 < pe2; b0 += SC_OCC) {
      const int32_t nb = (int32_t)((pe2 - b0) < SC_OCC ? (pe2 - b0) : SC_OCC);
      __syncthreads();  // previous batch done with the metadata
      for (int32_t e = tid; e < nb; e += SC_T) {
        const int32_t k = occ[b0 + e];
        const int32_t s = sl_of[k];
        const int32_t o = off[s];
        m_o[e] = o;
        m_ks[e] = off[s + 1] - o;
        m_q[e] = qoff[s] + (int64_t)(k - o) * qld[s];  // column j = k - o of Q_s
        m_sc[e] = scale ? scale[s] : 1.0;
      }
      __syncthreads();
      // indices clamped to the searchlight (ks >= 1: it holds column c's vertex),
      // so every load is valid and none sits behind a branch
      auto load = [&](int32_t e, ScRaw &x) {
        const int32_t o = m_o[e], ks = m_ks[e];
        const float *qc = q + m_q[e];
#pragma unroll
        for (int u = 0; u < SC_V; ++u) {
          const int32_t i = tid + u * SC_T;
          const int32_t ic = i < ks ? i : ks - 1;
          x.q[u] = qc[ic];
          x.w[u] = w ? w[o + ic] : 1.0;
          const int32_t r = slx_idx[o + ic];
          x.r[u] = i < ks ? r : -1;
        }
      };
      // contribution Q[i, j] * scale * w_i, in this order (as the oracle forms it)
      auto contrib = [&](float qv, double sc, double wv) {
        double t = (double)qv;
        if (scale) t = t * sc;
        return w ? t * wv : t;
      };
      for (int32_t e0 = 0; e0 < nb; e0 += SC_P) {
        // SC_P searchlights' loads in flight together, then their sums in order
        ScRaw x[SC_P];
#pragma unroll
        for (int p = 0; p < SC_P; ++p) load(e0 + p < nb ? e0 + p : nb - 1, x[p]);
#pragma unroll
        for (int p = 0; p < SC_P; ++p) {
          const int32_t e = e0 + p;
          if (e >= nb) break;  // uniform
          const double sc = m_sc[e];
#pragma unroll
          for (int u = 0; u < SC_V; ++u) {
            if (x[p].r[u] >= 0) {
              const int32_t pos = pos_of(x[p].r[u]);
              if (pos >= 0) acc[pos] += contrib(x[p].q[u], sc, x[p].w[u]);
            }
          }
          // searchlights larger than SC_V * SC_T: the rest, unpipelined
          const int32_t ks = m_ks[e];
          if (ks > SC_V * SC_T) {
            const int32_t o = m_o[e];
            const float *qc = q + m_q[e];
            for (int32_t i = tid + SC_V * SC_T; i < ks; i += SC_T) {
              const int32_t pos = pos_of(slx_idx[o + i]);
              if (pos >= 0) acc[pos] += contrib(qc[i], sc, w ? w[o + i] : 1.0);
            }
          }
          __syncthreads();
        }
      }
    }
    if (in_lds) {
      for (int32_t i = tid; i < L; i += SC_T) data[cb + i] = acc_l[i];
    }
    __syncthreads();
  }
}

extern "C" int hak_scatter(const int32_t *slx_idx, const int32_t *sl_off, const int32_t *sl_of,
                           const int64_t *occ_ptr, const int32_t *occ, int32_t ncols,
                           const int64_t *indptr, const int32_t *indices, double *data,
                           const float *q, const int64_t *qoff, const int32_t *qld,
                           const double *scale, const double *w,
                           int32_t s_begin, int32_t s_end, int32_t max_col, void *stream) {
  if (ncols <= 0) return 0;
  // LDS sized to the longest column (at most SC_LDS entries), hash at least twice that
  const int32_t cap = max_col < 1 ? 1 : (max_col > SC_LDS ? SC_LDS : max_col);
  int32_t hbits = 10;
  while ((1 << hbits) < 2 * cap) ++hbits;
  const size_t lds = (size_t)cap * sizeof(double) + (size_t)SC_OCC * (2 * sizeof(int64_t) + 2 * sizeof(int32_t)) +
                     ((size_t)1 << hbits) * sizeof(uint32_t);
  hipLaunchKernelGGL(k_scatter, dim3(ncols < 65536 ? ncols : 65536), dim3(SC_T), lds, (hipStream_t)stream,
                     slx_idx, sl_off, sl_of, occ_ptr, occ, ncols, indptr, indices, data, q, qoff, qld,
                     scale, w, s_begin, s_end, cap, hbits);
  LAUNCH_RET();
}

// One local transform into a CSC matrix (ha_sparse.c:128-142): thread per (i, j).
// Rows of one local transform are distinct vertices, so positions never collide.
__global__ void k_scatter_one(const int32_t *__restrict__ indptr, const int32_t *__restrict__ indices,
                              double *__restrict__ data, const int32_t *__restrict__ rows,
                              const int32_t *__restrict__ cols, int32_t nr, int32_t nc,
                              const double *__restrict__ T, const double *__restrict__ w) {
  int64_t n = (int64_t)nr * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int32_t i = (int32_t)(t / nc), j = (int32_t)(t % nc);
    int32_t col = cols[j];
    int32_t lo = indptr[col], hi = indptr[col + 1], e = hi;
    int32_t r = rows[i];
    while (lo < hi) {
      int32_t mid = (lo + hi) >> 1;
      if (indices[mid] < r) lo = mid + 1; else hi = mid;
    }
    if (lo < e && indices[lo] == r) {
      double val = w ? T[t] * w[i] : T[t] * 1.0;
      data[lo] += val;
    }
  }
}
extern "C" int hak_scatter_one(const int32_t *indptr, const int32_t *indices, double *data,
                               const int32_t *rows, const int32_t *cols, int32_t nr, int32_t nc,
                               const double *T, const double *w, void *stream) {
  int64_t n = (int64_t)nr * nc;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_scatter_one, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     indptr, indices, data, rows, cols, nr, nc, T, w);
  LAUNCH_RET();
}
