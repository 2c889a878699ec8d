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

__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
	for (int off = 32; off > 0; off >>= 1)
	{
		uint32_t o = __shfl_xor(v, off);
		v = o > v ? o : v;
	}
	return v;
}

// Workgroups of a kernel that ends in per-workgroup global atomics on a few
// counters: such atomics from every workgroup of a large grid serialize, so
// the kernel loops over its elements on at most this many workgroups.
constexpr unsigned kReduceGroups = 1024;

static inline unsigned reduce_grid(int64_t n, int block)
{
	int64_t g = (n + block - 1) / block;
	return (unsigned)(g < 1 ? 1 : (g > (int64_t)kReduceGroups ? (int64_t)kReduceGroups : g));
}

// *p += sum of v over the wave.
__device__ __forceinline__ void wave_add(uint32_t* p, uint32_t v)
{
	uint32_t s = wave_sum(v);
	if (__lane_id() == 0 && s != 0u)
		atomicAdd(p, s);
}

// *p += number of calling lanes. Unlike the helpers above it may be called
// from divergent code: only the wave's active lanes take part.
__device__ __forceinline__ void wave_inc(uint32_t* p)
{
	uint64_t m = __ballot(1);
	if (__lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1))
		atomicAdd(p, (uint32_t)__popcll(m));
}

// ctr[idx[k]] += sum of v[k] over the workgroup, for k < N: one atomic per
// counter and workgroup, for kernels in which every wave adds to the same few
// counters. Every lane of the workgroup calls it (it holds a barrier); at most
// 16 waves per workgroup.
template <int N>
__device__ __forceinline__ void block_add(uint32_t* ctr, const int (&idx)[N], const uint32_t (&v)[N])
{
	__shared__ uint32_t part[N][16];
	const uint32_t wv = threadIdx.x >> 6;
#pragma unroll
	for (int k = 0; k < N; ++k)
	{
		uint32_t s = wave_sum(v[k]);
		if (__lane_id() == 0)
			part[k][wv] = s;
	}
	__syncthreads();
	if (threadIdx.x < (uint32_t)N)
	{
		const uint32_t nw = (blockDim.x + 63) >> 6;
		uint32_t t = 0;
		for (uint32_t w = 0; w < nw; ++w)
			t += part[threadIdx.x][w];
		int k = (int)threadIdx.x;
		int at = idx[0];
#pragma unroll
		for (int j = 1; j < N; ++j)
			at = k == j ? idx[j] : at;
		if (t != 0u)
			atomicAdd(&ctr[at], t);
	}
}

// Rounds of the keyed helpers below: each round sums the lanes of one key
// into one atomic; the lanes left after them (keys shared by few lanes, as
// the one-body islands of a scattered world, which would take a round per
// lane) issue their own atomics together.
constexpr int kKeyedRounds = 4;

// ctr[key] += v for every lane whose key is not ~0u.
__device__ __forceinline__ void wave_add_keyed(uint32_t* ctr, uint32_t key, uint32_t v)
{
	uint64_t todo = __ballot(key != ~0u);
	for (int round = 0; round < kKeyedRounds && todo; ++round)
	{
		int leader = __ffsll((unsigned long long)todo) - 1;
		uint32_t lk = __shfl(key, leader);
		bool mine = key == lk;
		uint32_t s = wave_sum(mine ? v : 0u);
		if ((int)__lane_id() == leader)
			atomicAdd(&ctr[lk], s);
		todo &= ~__ballot(mine);
	}
	if ((todo >> __lane_id()) & 1ull)
		atomicAdd(&ctr[key], v);
}

// ctr[key] = min(ctr[key], v) for every lane whose key is not ~0u.
__device__ __forceinline__ void wave_min_keyed(uint32_t* ctr, uint32_t key, uint32_t v)
{
	uint64_t todo = __ballot(key != ~0u);
	for (int round = 0; round < kKeyedRounds && todo; ++round)
	{
		int leader = __ffsll((unsigned long long)todo) - 1;
		uint32_t lk = __shfl(key, leader);
		bool mine = key == lk;
		uint32_t m = wave_min(mine ? v : 0xFFFFFFFFu);
		if ((int)__lane_id() == leader)
			atomicMin(&ctr[lk], m);
		todo &= ~__ballot(mine);
	}
	if ((todo >> __lane_id()) & 1ull)
		atomicMin(&ctr[key], v);
}

} // namespace mx
