// cs_prims.hpp — device-wide primitives written for gfx950 (64-wide wavefronts):
// exclusive scan of uint32 and a stable LSD radix sort of (uint64 key, uint32 value).
#pragma once
#include "cs_common.hpp"

namespace cs {

struct Scratch {
    DBuf a, b, c, d;
};

// out[i] = sum(in[0..i)), out[n] = total.  `out` must hold n+1 entries; in == out is allowed.
void scan_exclusive(const uint32_t* in, uint32_t* out, int n, Scratch& ws, hipStream_t s);

// Stable sort of n (key, value) pairs by key bits [bit0, bit1).  The sorted data
// ends in (keys, vals); (keys_tmp, vals_tmp) are n-element temporaries.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp, int n, int bit0,
                      int bit1, Scratch& ws, hipStream_t s);

// Fill helpers usable inside captured graphs (no hipMemsetAsync).
void fill_u32(uint32_t* p, uint32_t v, long n, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t v, long n, hipStream_t s);
void fill_f32(float* p, float v, long n, hipStream_t s);
void iota_u32(uint32_t* p, long n, hipStream_t s);

}  // namespace cs
