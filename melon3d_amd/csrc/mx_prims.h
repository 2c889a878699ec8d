// Device-wide primitives used by the step pipeline: exclusive scan, stable
// LSD radix sort of (u64 key, u32 value) pairs, and a u64-keyed hash set.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mx
{

// Scratch arena handed to the primitives (grown by the caller).
struct PrimScratch
{
	void* buf = nullptr;
	size_t bytes = 0;
};

// Bytes of scratch needed by scan_exclusive for n elements.
size_t scan_scratch_bytes(int64_t n);
// out[i] = sum(in[0..i)); *total (device pointer, may be null) = sum(in).
// in and out may alias.
void scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, void* scratch, hipStream_t s);

// Bytes of scratch needed by radix_sort_pairs for n elements.
size_t sort_scratch_bytes(int64_t n);
// Stable sort of (keys, vals) by the low `bits` bits of keys, ascending.
// Results end in keys/vals; keysAlt/valsAlt are ping-pong buffers of n.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keysAlt, uint32_t* valsAlt, int64_t n, int bits,
					  void* scratch, hipStream_t s);

} // namespace mx
