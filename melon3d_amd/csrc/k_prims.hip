// Scan and stable radix sort for gfx950: 256-thread workgroups (4 waves of
// 64 lanes). The sort ranks each element inside its tile with wave ballots
// (one ballot per digit bit) plus per-wave digit counts in LDS, so the
// scatter is stable without per-element atomics.
#include "mx_prims.h"

#include <stdio.h>

#include <algorithm>
#include <atomic>

namespace mx
{

constexpr int kScanBlock = 1024; // elements per scan workgroup (4 per thread)
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* lds, uint32_t* blockTotal)
{
	// lds: 256 entries
	const int t = threadIdx.x;
	lds[t] = v;
	__syncthreads();
	for (int off = 1; off < 256; off <<= 1)
	{
		uint32_t add = t >= off ? lds[t - off] : 0u;
		__syncthreads();
		lds[t] += add;
		__syncthreads();
	}
	uint32_t incl = lds[t];
	*blockTotal = lds[255];
	__syncthreads();
	return incl - v;
}

// ---------------------------------------------------------------- scan
//
// Single-pass exclusive scan with decoupled look-back: each workgroup scans
// its 2048-element tile, publishes the tile sum, then adds the prefix of the
// tiles before it, read from their published words (an inclusive prefix ends
// the look-back early). Words are tagged with the call's epoch, so the
// status array never needs clearing between calls: it is zeroed once when
// allocated and holds nothing but status words (any other data left in it
// could carry the current tag and be read as a published sum). A workgroup only waits on lower-numbered ones,
// which publish their sums without waiting, so the look-back always ends.

constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;
constexpr uint64_t kFlagAgg = 1ull, kFlagPrefix = 2ull;

__device__ __forceinline__ uint64_t pack_status(uint32_t epoch, uint64_t flag, uint32_t v)
{
	return ((uint64_t)epoch << 34) | (flag << 32) | v;
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
	const int lane = __lane_id();
	for (int off = 1; off < 64; off <<= 1)
	{
		uint32_t o = __shfl_up(v, off);
		if (lane >= off)
			v += o;
	}
	return v;
}

__global__ void __launch_bounds__(256) k_scan_onepass(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total,
													  uint64_t* status, uint32_t epoch)
{
	__shared__ uint32_t waveSum[4];
	__shared__ uint32_t tilePrefix;
	const int t = threadIdx.x;
	const int lane = t & 63, wave = t >> 6;
	const int64_t tile = blockIdx.x;
	const int64_t base = tile * kScanTile + (int64_t)t * kScanItems;
	uint32_t v[kScanItems];
	uint32_t sum = 0;
	const bool vecIn = ((uintptr_t)in & 15u) == 0u;
	const bool vecOut = ((uintptr_t)out & 15u) == 0u;
	if (vecIn && base + kScanItems <= n)
	{
		const uint4* p = (const uint4*)(in + base);
		uint4 a = p[0], b = p[1];
		v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
		v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
	}
	else
	{
#pragma unroll
		for (int k = 0; k < kScanItems; ++k)
			v[k] = base + k < n ? in[base + k] : 0u;
	}
#pragma unroll
	for (int k = 0; k < kScanItems; ++k)
		sum += v[k];
	uint32_t incl = wave_inclusive_sum(sum);
	if (lane == 63)
		waveSum[wave] = incl;
	__syncthreads();
	uint32_t waveOff = 0, agg = 0;
	for (int w = 0; w < 4; ++w)
	{
		waveOff += w < wave ? waveSum[w] : 0u;
		agg += waveSum[w];
	}
	// publish, then look back (wave 0)
	if (wave == 0)
	{
		uint32_t prefix = 0;
		if (tile == 0)
		{
			if (lane == 0)
				__hip_atomic_store(&status[0], pack_status(epoch, kFlagPrefix, agg), __ATOMIC_RELAXED,
								   __HIP_MEMORY_SCOPE_AGENT);
		}
		else
		{
			if (lane == 0)
				__hip_atomic_store(&status[tile], pack_status(epoch, kFlagAgg, agg), __ATOMIC_RELAXED,
								   __HIP_MEMORY_SCOPE_AGENT);
			int64_t j = tile - 1; // nearest tile not yet consumed
			for (;;)
			{
				// lane l reads tile j - l; before tile 0 a zero prefix stands in
				const int64_t idx = j - lane;
				uint64_t word;
				uint64_t preMask;
				int first;
				for (;;)
				{
					word = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
									: pack_status(epoch, kFlagPrefix, 0u);
					uint64_t flag = (word >> 32) & 3ull;
					bool ready = (uint32_t)(word >> 34) == epoch && flag != 0ull;
					uint64_t readyMask = __ballot(ready);
					preMask = __ballot(ready && flag == kFlagPrefix);
					first = preMask ? __ffsll((unsigned long long)preMask) - 1 : 64;
					// lanes 0..first (all 64 without a prefix) must be ready
					uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
					if ((readyMask & need) == need)
						break;
				}
				uint32_t mine = lane <= first ? (uint32_t)word : 0u;
				for (int off = 32; off > 0; off >>= 1)
					mine += __shfl_xor(mine, off);
				prefix += mine;
				if (first < 64)
					break;
				j -= 64;
			}
			if (lane == 0)
				__hip_atomic_store(&status[tile], pack_status(epoch, kFlagPrefix, prefix + agg), __ATOMIC_RELAXED,
								   __HIP_MEMORY_SCOPE_AGENT);
		}
		if (lane == 0)
		{
			tilePrefix = prefix;
			if (total != nullptr && (tile + 1) * kScanTile >= n)
				*total = prefix + agg;
		}
	}
	__syncthreads();
	uint32_t run = tilePrefix + waveOff + (incl - sum);
	if (vecOut && base + kScanItems <= n)
	{
		uint32_t o[kScanItems];
#pragma unroll
		for (int k = 0; k < kScanItems; ++k)
		{
			o[k] = run;
			run += v[k];
		}
		uint4* q = (uint4*)(out + base);
		q[0] = make_uint4(o[0], o[1], o[2], o[3]);
		q[1] = make_uint4(o[4], o[5], o[6], o[7]);
	}
	else
	{
#pragma unroll
		for (int k = 0; k < kScanItems; ++k)
		{
			if (base + k < n)
				out[base + k] = run;
			run += v[k];
		}
	}
}

static uint32_t next_scan_epoch()
{
	static std::atomic<uint32_t> epoch{ 0 };
	return (epoch.fetch_add(1, std::memory_order_relaxed) + 1u) & 0x3FFFFFFFu;
}

int64_t scan_status_words(int64_t n)
{
	return (n + kScanTile - 1) / kScanTile + 1;
}

void scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint64_t* status, hipStream_t s)
{
	if (n <= 0)
	{
		if (total != nullptr)
		{
			hipMemsetAsync(total, 0, sizeof(uint32_t), s);
		}
		return;
	}
	int64_t tiles = (n + kScanTile - 1) / kScanTile;
	// in and out may alias: each workgroup reads its whole tile before it
	// writes it, and reads no other tile's input
	k_scan_onepass<<<(unsigned)tiles, 256, 0, s>>>(in, out, n, total, status, next_scan_epoch());
}

// ------------------------------------------------------------ radix sort
//
// LSD radix sort of (key, value) pairs, 8-bit digits, one launch per digit
// pass after one histogram launch for all passes:
// - k_radix_digits counts every pass's digits over all keys (LDS tables,
//   one global atomic per digit and workgroup);
// - k_radix_pass sorts one digit. A workgroup takes a tile of kSortTile
//   consecutive pairs (kSortItems per lane, read wave-striped: item j of lane
//   l of wave w is element w * 64 * kSortItems + j * 64 + l of the tile, so
//   each load instruction reads 64 consecutive keys), ranks them stably in
//   that order (per wave: match masks from 8 ballots and a running per-digit
//   count in LDS; then the waves' counts are offset per digit), finds each
//   digit's global start for the tile with a decoupled look-back over the
//   tiles before it (one lane per digit; status words tagged with the call's
//   epoch as in the scan), stages the tile in LDS in sorted order and writes
//   it out from there, so consecutive lanes write consecutive addresses of
//   each digit's run.
// Stable: within a tile the rank follows the element order, and tiles take
// their digit offsets in tile order. A tile only waits on lower-numbered
// tiles, which publish their counts before they wait (see the scan).

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems; // 4096 pairs per tile
constexpr int kMaxPasses = 8;

__device__ __forceinline__ uint32_t digit_of(uint64_t k, int shift) { return (uint32_t)(k >> shift) & (kRadix - 1); }

__global__ void k_radix_zero(uint32_t* p, int n)
{
	int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n)
		p[i] = 0u;
}

// hist[p * 256 + d]: keys whose digit p is d
__global__ void __launch_bounds__(256) k_radix_digits(const uint64_t* keys, int64_t n, int passes, uint32_t* hist)
{
	__shared__ uint32_t cnt[kMaxPasses * kRadix];
	for (int k = threadIdx.x; k < passes * kRadix; k += 256)
		cnt[k] = 0u;
	__syncthreads();
	for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
	{
		uint64_t key = keys[i];
		for (int p = 0; p < passes; ++p)
			atomicAdd(&cnt[p * kRadix + digit_of(key, p * kRadixBits)], 1u);
	}
	__syncthreads();
	for (int k = threadIdx.x; k < passes * kRadix; k += 256)
		if (cnt[k] != 0u)
			atomicAdd(&hist[k], cnt[k]);
}

__global__ void __launch_bounds__(kSortThreads) k_radix_pass(const uint64_t* keys, const uint32_t* vals,
															uint64_t* keysOut, uint32_t* valsOut, int64_t n, int shift,
															const uint32_t* digitCount, uint64_t* status,
															uint32_t epoch)
{
	__shared__ uint64_t sKey[kSortTile];
	__shared__ uint32_t sVal[kSortTile];
	__shared__ uint32_t waveCnt[4][kRadix];
	__shared__ uint32_t scanTmp[kRadix];
	__shared__ uint32_t tileStart[kRadix]; // tile-local start of each digit
	__shared__ int64_t outBase[kRadix];    // global position of the tile's first key of each digit, minus tileStart
	const int t = threadIdx.x;
	const int lane = t & 63, wave = t >> 6;
	const int64_t tile = blockIdx.x;
	const int64_t tileBase = tile * kSortTile;
	for (int k = t; k < 4 * kRadix; k += kSortThreads)
		(&waveCnt[0][0])[k] = 0u;
	__syncthreads();

	uint64_t key[kSortItems];
	uint32_t val[kSortItems];
	uint32_t rank[kSortItems];
	const int64_t waveBase = tileBase + (int64_t)wave * 64 * kSortItems;
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		int64_t i = waveBase + j * 64 + lane;
		bool valid = i < n;
		key[j] = valid ? keys[i] : 0ull;
		val[j] = valid ? vals[i] : 0u;
	}
	const uint64_t lowerMask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		int64_t i = waveBase + j * 64 + lane;
		bool valid = i < n;
		uint32_t d = digit_of(key[j], shift);
		uint64_t same = __ballot(valid);
#pragma unroll
		for (int b = 0; b < kRadixBits; ++b)
		{
			bool bit = (d >> b) & 1u;
			uint64_t m = __ballot(bit);
			same &= bit ? m : ~m;
		}
		uint32_t before = valid ? waveCnt[wave][d] : 0u;
		uint32_t below = (uint32_t)__popcll(same & lowerMask);
		rank[j] = before + below;
		// the group's lowest lane advances the wave's count of the digit
		if (valid && below == 0u)
			waveCnt[wave][d] = before + (uint32_t)__popcll(same);
		__builtin_amdgcn_wave_barrier();
	}
	__syncthreads();
	// lane t owns digit t: the waves' offsets, the tile's count
	uint32_t tileCount = 0;
	{
		uint32_t c[4];
#pragma unroll
		for (int w = 0; w < 4; ++w)
			c[w] = waveCnt[w][t];
#pragma unroll
		for (int w = 0; w < 4; ++w)
		{
			waveCnt[w][t] = tileCount;
			tileCount += c[w];
		}
	}
	uint32_t blockTotal;
	uint32_t start = block_exclusive_sum(tileCount, scanTmp, &blockTotal);
	tileStart[t] = start;
	// the digit's global start: keys of smaller digits, then this digit's
	// keys in the tiles before this one
	uint32_t allBefore = block_exclusive_sum(digitCount[t], scanTmp, &blockTotal);
	uint64_t* st = status + tile * kRadix;
	uint32_t prefix = 0;
	if (tile == 0)
	{
		__hip_atomic_store(&st[t], pack_status(epoch, kFlagPrefix, tileCount), __ATOMIC_RELAXED,
						   __HIP_MEMORY_SCOPE_AGENT);
	}
	else
	{
		__hip_atomic_store(&st[t], pack_status(epoch, kFlagAgg, tileCount), __ATOMIC_RELAXED,
						   __HIP_MEMORY_SCOPE_AGENT);
		for (int64_t j = tile - 1; j >= 0; --j)
		{
			uint64_t word;
			for (;;)
			{
				word = __hip_atomic_load(&status[j * kRadix + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				if ((uint32_t)(word >> 34) == epoch && ((word >> 32) & 3ull) != 0ull)
					break;
				__builtin_amdgcn_s_sleep(1);
			}
			prefix += (uint32_t)word;
			if (((word >> 32) & 3ull) == kFlagPrefix)
				break;
		}
		__hip_atomic_store(&st[t], pack_status(epoch, kFlagPrefix, prefix + tileCount), __ATOMIC_RELAXED,
						   __HIP_MEMORY_SCOPE_AGENT);
	}
	outBase[t] = (int64_t)allBefore + (int64_t)prefix - (int64_t)start;
	__syncthreads();
	// stage in tile order of the digit, then write out
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		int64_t i = waveBase + j * 64 + lane;
		if (i < n)
		{
			uint32_t d = digit_of(key[j], shift);
			uint32_t at = tileStart[d] + waveCnt[wave][d] + rank[j];
			sKey[at] = key[j];
			sVal[at] = val[j];
		}
	}
	__syncthreads();
	const int tileN = (int)(n - tileBase < kSortTile ? n - tileBase : kSortTile);
	for (int at = t; at < tileN; at += kSortThreads)
	{
		uint64_t k = sKey[at];
		int64_t dst = outBase[digit_of(k, shift)] + at;
		keysOut[dst] = k;
		valsOut[dst] = sVal[at];
	}
}

size_t sort_scratch_bytes(int64_t)
{
	return (size_t)kMaxPasses * kRadix * sizeof(uint32_t);
}

int64_t sort_status_words(int64_t n)
{
	int64_t tiles = (n + kSortTile - 1) / kSortTile;
	return tiles * kRadix + 1;
}

static thread_local const SortTimer* t_sortTimer = nullptr;

void set_sort_timer(const SortTimer* t) { t_sortTimer = t; }

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keysAlt, uint32_t* valsAlt, int64_t n, int bits,
					  void* scratch, uint64_t* status, hipStream_t s)
{
	if (n <= 1 || bits <= 0)
	{
		return;
	}
	const int passes = std::min((bits + kRadixBits - 1) / kRadixBits, kMaxPasses);
	const int64_t tiles = (n + kSortTile - 1) / kSortTile;
	uint32_t* hist = (uint32_t*)scratch;
	// algorithmic bytes: the keys read once for the digit counts (8 B), then
	// every (key, value) pair read and written once per pass (24 B)
	const SortTimer* timer = t_sortTimer;
	if (timer != nullptr)
		timer->begin(timer->ctx, 8.0 * (double)n + 24.0 * (double)n * passes);
	k_radix_zero<<<(passes * kRadix + 255) / 256, 256, 0, s>>>(hist, passes * kRadix);
	const int64_t histBlocks = std::min<int64_t>((n + 4095) / 4096, 1024);
	k_radix_digits<<<(unsigned)histBlocks, 256, 0, s>>>(keys, n, passes, hist);
	uint64_t* kin = keys;
	uint32_t* vin = vals;
	uint64_t* kout = keysAlt;
	uint32_t* vout = valsAlt;
	for (int p = 0; p < passes; ++p)
	{
		k_radix_pass<<<(unsigned)tiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, p * kRadixBits,
															  hist + p * kRadix, status, next_scan_epoch());
		std::swap(kin, kout);
		std::swap(vin, vout);
	}
	if (kin != keys)
	{
		hipMemcpyAsync(keys, kin, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
		hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
	}
	if (timer != nullptr)
		timer->end(timer->ctx);
}

} // namespace mx
