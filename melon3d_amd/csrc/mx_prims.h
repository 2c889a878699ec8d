// Device-wide primitives used by the step pipeline: exclusive scan, stable
// LSD radix sort of (u64 key, u32 value) pairs, and a u64-keyed hash set.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mx
{

// Look-back status words scan_exclusive needs for n elements. The status
// storage belongs to the scans alone: zeroed once when allocated and never
// written by anything else (its words carry a per-call epoch tag, so a
// scan tells the words of earlier scans apart, but not arbitrary data).
int64_t scan_status_words(int64_t n);
// out[i] = sum(in[0..i)); *total (device pointer, may be null) = sum(in).
// in and out may alias.
void scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint64_t* status, hipStream_t s);
// Two or three independent exclusive scans of n elements each in one launch
// (status holds scan_status_words(n) words; in and out of an array may alias).
void scan_exclusive2(const uint32_t* in0, uint32_t* out0, uint32_t* total0, const uint32_t* in1, uint32_t* out1,
					 uint32_t* total1, int64_t n, uint64_t* status, hipStream_t s);
void scan_exclusive3(const uint32_t* in0, uint32_t* out0, uint32_t* total0, const uint32_t* in1, uint32_t* out1,
					 uint32_t* total1, const uint32_t* in2, uint32_t* out2, uint32_t* total2, int64_t n,
					 uint64_t* status, hipStream_t s);

// Bytes of histogram scratch and status words radix_sort_pairs needs for n
// elements.
size_t sort_scratch_bytes(int64_t n);
int64_t sort_status_words(int64_t n);
// Stable sort of (keys, vals) by the low `bits` bits of keys, ascending.
// Results end in keys/vals; keysAlt/valsAlt are ping-pong buffers of n.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keysAlt, uint32_t* valsAlt, int64_t n, int bits,
					  void* scratch, uint64_t* status, hipStream_t s);
// The same sort without the copy back after an odd number of passes: true
// when the results are in keysAlt/valsAlt.
bool radix_sort_pairs_in_place_or_alt(uint64_t* keys, uint32_t* vals, uint64_t* keysAlt, uint32_t* valsAlt, int64_t n,
									  int bits, void* scratch, uint64_t* status, hipStream_t s);

// Kernel timing of the sorts issued by one thread: while a timer is set,
// radix_sort_pairs calls begin(ctx, bytes) before its passes, with the
// sort's algorithmic bytes, and end(ctx) after them.
struct SortTimer
{
	void* ctx;
	void (*begin)(void* ctx, double bytes);
	void (*end)(void* ctx);
};
void set_sort_timer(const SortTimer* t);

} // namespace mx
