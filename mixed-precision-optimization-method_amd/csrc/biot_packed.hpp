// Per-cell packed blocks in 16-byte groups (device side).
//
// A block holds N values of type T per cell.  The values of a cell are split into 16-byte
// groups, stored group-major: group g (values g*G .. g*G+G-1, G = 16 / sizeof(T)) is an array
// of n 16-byte records at byte g*n*16, and the last N mod G values form a tail array of
// (N mod G)*sizeof(T)-byte records after the full groups.  A wavefront's load or store of one
// group is then one contiguous run (a dwordx4, or the tail's dwordx3/x2/x1, per lane).
// Addresses are 64-bit: a block may span more than 4 GiB.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace biot {
namespace packed {

typedef uint32_t dw4_t __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t dw3_t __attribute__((ext_vector_type(3), aligned(4)));
typedef uint32_t dw2_t __attribute__((ext_vector_type(2), aligned(4)));

template <int ND, bool NT>
__device__ __forceinline__ void ld_rec(const char *p, uint32_t *d)
{
  if constexpr (ND == 4) {
    const dw4_t *q = reinterpret_cast<const dw4_t *>(p);
    const dw4_t v = NT ? __builtin_nontemporal_load(q) : *q;
    d[0] = v.x, d[1] = v.y, d[2] = v.z, d[3] = v.w;
  } else if constexpr (ND == 3) {
    const dw3_t *q = reinterpret_cast<const dw3_t *>(p);
    const dw3_t v = NT ? __builtin_nontemporal_load(q) : *q;
    d[0] = v.x, d[1] = v.y, d[2] = v.z;
  } else if constexpr (ND == 2) {
    const dw2_t *q = reinterpret_cast<const dw2_t *>(p);
    const dw2_t v = NT ? __builtin_nontemporal_load(q) : *q;
    d[0] = v.x, d[1] = v.y;
  } else if constexpr (ND == 1) {
    const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
    d[0] = NT ? __builtin_nontemporal_load(q) : *q;
  }
}

template <int ND>
__device__ __forceinline__ void st_rec(char *p, const uint32_t *d)
{
  if constexpr (ND == 4) {
    dw4_t v;
    v.x = d[0], v.y = d[1], v.z = d[2], v.w = d[3];
    *reinterpret_cast<dw4_t *>(p) = v;
  } else if constexpr (ND == 3) {
    dw3_t v;
    v.x = d[0], v.y = d[1], v.z = d[2];
    *reinterpret_cast<dw3_t *>(p) = v;
  } else if constexpr (ND == 2) {
    dw2_t v;
    v.x = d[0], v.y = d[1];
    *reinterpret_cast<dw2_t *>(p) = v;
  } else if constexpr (ND == 1) {
    *reinterpret_cast<uint32_t *>(p) = d[0];
  }
}

// dwords of a block of N values of type T
template <typename T, int N>
constexpr int block_dwords()
{
  return (N * (int)sizeof(T) + 3) / 4;
}

// the raw dwords of cell c (n cells in the block)
template <typename T, int N, bool NT>
__device__ __forceinline__ void load(const void *base, uint64_t n, uint32_t c, uint32_t *d)
{
  constexpr int G = 16 / (int)sizeof(T);
  constexpr int NG = N / G, TB = (N % G) * (int)sizeof(T);
  static_assert(TB % 4 == 0, "tail records must be whole dwords");
  const char *b = reinterpret_cast<const char *>(base);
#pragma unroll
  for (int g = 0; g < NG; ++g) ld_rec<4, NT>(b + (uint64_t)g * n * 16u + (uint64_t)c * 16u, d + 4 * g);
  if constexpr (TB > 0) ld_rec<TB / 4, NT>(b + (uint64_t)NG * n * 16u + (uint64_t)c * TB, d + 4 * NG);
}

template <typename T, int N>
__device__ __forceinline__ void store(void *base, uint64_t n, uint32_t c, const uint32_t *d)
{
  constexpr int G = 16 / (int)sizeof(T);
  constexpr int NG = N / G, TB = (N % G) * (int)sizeof(T);
  static_assert(TB % 4 == 0, "tail records must be whole dwords");
  char *b = reinterpret_cast<char *>(base);
#pragma unroll
  for (int g = 0; g < NG; ++g) st_rec<4>(b + (uint64_t)g * n * 16u + (uint64_t)c * 16u, d + 4 * g);
  if constexpr (TB > 0) st_rec<TB / 4>(b + (uint64_t)NG * n * 16u + (uint64_t)c * TB, d + 4 * NG);
}

// value e of a block held as dwords
template <typename T>
__device__ __forceinline__ T value(const uint32_t *d, int e)
{
  if constexpr (sizeof(T) == 2) {
    return __builtin_bit_cast(T, (uint16_t)(d[e >> 1] >> (16 * (e & 1))));
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, d[e]);
  } else {
    return __builtin_bit_cast(T, (uint64_t)d[2 * e] | ((uint64_t)d[2 * e + 1] << 32));
  }
}

}  // namespace packed
}  // namespace biot
