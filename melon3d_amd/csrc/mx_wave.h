// Wave-level aggregation of global atomics (device code only). Atomics on
// one address from many lanes serialize in the memory system; these helpers
// combine a wave's contributions first, so a wave issues one atomic per
// distinct address. All 64 lanes of the wave must call them together.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mx
{

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_xor(v, off);
	return v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
	for (int off = 32; off > 0; off >>= 1)
	{
		uint32_t o = __shfl_xor(v, off);
		v = o < v ? o : v;
	}
	return v;
}

// *p += sum of v over the wave.
__device__ __forceinline__ void wave_add(uint32_t* p, uint32_t v)
{
	uint32_t s = wave_sum(v);
	if (__lane_id() == 0 && s != 0u)
		atomicAdd(p, s);
}

// ctr[key] += v for every lane whose key is not ~0u.
__device__ __forceinline__ void wave_add_keyed(uint32_t* ctr, uint32_t key, uint32_t v)
{
	uint64_t todo = __ballot(key != ~0u);
	while (todo)
	{
		int leader = __ffsll((unsigned long long)todo) - 1;
		uint32_t lk = __shfl(key, leader);
		bool mine = key == lk;
		uint32_t s = wave_sum(mine ? v : 0u);
		if ((int)__lane_id() == leader)
			atomicAdd(&ctr[lk], s);
		todo &= ~__ballot(mine);
	}
}

// ctr[key] = min(ctr[key], v) for every lane whose key is not ~0u.
__device__ __forceinline__ void wave_min_keyed(uint32_t* ctr, uint32_t key, uint32_t v)
{
	uint64_t todo = __ballot(key != ~0u);
	while (todo)
	{
		int leader = __ffsll((unsigned long long)todo) - 1;
		uint32_t lk = __shfl(key, leader);
		bool mine = key == lk;
		uint32_t m = wave_min(mine ? v : 0xFFFFFFFFu);
		if ((int)__lane_id() == leader)
			atomicMin(&ctr[lk], m);
		todo &= ~__ballot(mine);
	}
}

} // namespace mx
