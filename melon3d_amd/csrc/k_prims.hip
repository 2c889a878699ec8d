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

// K independent exclusive scans of n elements in one launch (K <= 3): the
// same single-pass scan per array, wave k looking back over array k's status
// words (status + k * tiles). Saves the launches of scans that follow one
// flag kernel.
template <int K>
struct ScanArrays
{
	const uint32_t* in[K];
	uint32_t* out[K];
	uint32_t* total[K];
};

template <int K>
__global__ void __launch_bounds__(256) k_scan_multi(ScanArrays<K> A, int64_t n, uint64_t* status, int64_t tiles,
													uint32_t epoch)
{
	static_assert(K >= 1 && K <= 3, "one look-back wave per array");
	__shared__ uint32_t waveSum[K][4];
	__shared__ uint32_t tilePrefix[K];
	const int t = threadIdx.x;
	const int lane = t & 63, wave = t >> 6;
	const int64_t tile = blockIdx.x;
	const int64_t base = tile * kScanTile + (int64_t)t * kScanItems;
	uint32_t v[K][kScanItems];
	uint32_t sum[K], incl[K];
#pragma unroll
	for (int a = 0; a < K; ++a)
	{
		sum[a] = 0u;
#pragma unroll
		for (int k = 0; k < kScanItems; ++k)
		{
			v[a][k] = base + k < n ? A.in[a][base + k] : 0u;
			sum[a] += v[a][k];
		}
		incl[a] = wave_inclusive_sum(sum[a]);
		if (lane == 63)
			waveSum[a][wave] = incl[a];
	}
	__syncthreads();
	uint32_t waveOff[K], agg[K];
#pragma unroll
	for (int a = 0; a < K; ++a)
	{
		waveOff[a] = 0u;
		agg[a] = 0u;
		for (int w = 0; w < 4; ++w)
		{
			waveOff[a] += w < wave ? waveSum[a][w] : 0u;
			agg[a] += waveSum[a][w];
		}
	}
	if (wave < K)
	{
		// wave k publishes and looks back for array k
		uint32_t myAgg = agg[0];
#pragma unroll
		for (int a = 1; a < K; ++a)
			myAgg = wave == a ? agg[a] : myAgg;
		uint64_t* st = status + (int64_t)wave * tiles;
		uint32_t prefix = 0;
		if (tile == 0)
		{
			if (lane == 0)
				__hip_atomic_store(&st[0], pack_status(epoch, kFlagPrefix, myAgg), __ATOMIC_RELAXED,
								   __HIP_MEMORY_SCOPE_AGENT);
		}
		else
		{
			if (lane == 0)
				__hip_atomic_store(&st[tile], pack_status(epoch, kFlagAgg, myAgg), __ATOMIC_RELAXED,
								   __HIP_MEMORY_SCOPE_AGENT);
			int64_t j = tile - 1;
			for (;;)
			{
				const int64_t idx = j - lane;
				uint64_t word;
				uint64_t preMask;
				int first;
				for (;;)
				{
					word = idx >= 0 ? __hip_atomic_load(&st[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
									: pack_status(epoch, kFlagPrefix, 0u);
					uint64_t flag = (word >> 32) & 3ull;
					bool ready = (uint32_t)(word >> 34) == epoch && flag != 0ull;
					uint64_t readyMask = __ballot(ready);
					preMask = __ballot(ready && flag == kFlagPrefix);
					first = preMask ? __ffsll((unsigned long long)preMask) - 1 : 64;
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
				__hip_atomic_store(&st[tile], pack_status(epoch, kFlagPrefix, prefix + myAgg), __ATOMIC_RELAXED,
								   __HIP_MEMORY_SCOPE_AGENT);
		}
		if (lane == 0)
		{
			tilePrefix[wave] = prefix;
			uint32_t* tot = A.total[0];
#pragma unroll
			for (int a = 1; a < K; ++a)
				tot = wave == a ? A.total[a] : tot;
			if (tot != nullptr && (tile + 1) * kScanTile >= n)
				*tot = prefix + myAgg;
		}
	}
	__syncthreads();
#pragma unroll
	for (int a = 0; a < K; ++a)
	{
		uint32_t run = tilePrefix[a] + waveOff[a] + (incl[a] - sum[a]);
#pragma unroll
		for (int k = 0; k < kScanItems; ++k)
		{
			if (base + k < n)
				A.out[a][base + k] = run;
			run += v[a][k];
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
	// room for the three arrays of scan_exclusive3
	return 3 * ((n + kScanTile - 1) / kScanTile) + 1;
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

template <int K>
static void scan_exclusive_k(const uint32_t* const* in, uint32_t* const* out, int64_t n, uint32_t* const* total,
							 uint64_t* status, hipStream_t s)
{
	if (n <= 0)
	{
		for (int a = 0; a < K; ++a)
			if (total[a] != nullptr)
				hipMemsetAsync(total[a], 0, sizeof(uint32_t), s);
		return;
	}
	ScanArrays<K> A;
	for (int a = 0; a < K; ++a)
	{
		A.in[a] = in[a];
		A.out[a] = out[a];
		A.total[a] = total[a];
	}
	int64_t tiles = (n + kScanTile - 1) / kScanTile;
	k_scan_multi<K><<<(unsigned)tiles, 256, 0, s>>>(A, n, status, tiles, next_scan_epoch());
}

void scan_exclusive2(const uint32_t* in0, uint32_t* out0, uint32_t* total0, const uint32_t* in1, uint32_t* out1,
					 uint32_t* total1, int64_t n, uint64_t* status, hipStream_t s)
{
	const uint32_t* in[2] = { in0, in1 };
	uint32_t* out[2] = { out0, out1 };
	uint32_t* total[2] = { total0, total1 };
	scan_exclusive_k<2>(in, out, n, total, status, s);
}

void scan_exclusive3(const uint32_t* in0, uint32_t* out0, uint32_t* total0, const uint32_t* in1, uint32_t* out1,
					 uint32_t* total1, const uint32_t* in2, uint32_t* out2, uint32_t* total2, int64_t n,
					 uint64_t* status, hipStream_t s)
{
	const uint32_t* in[3] = { in0, in1, in2 };
	uint32_t* out[3] = { out0, out1, out2 };
	uint32_t* total[3] = { total0, total1, total2 };
	scan_exclusive_k<3>(in, out, n, total, status, s);
}

// ------------------------------------------------------------ radix sort
//
// LSD radix sort of (key, value) pairs, 8-bit digits, on tiles of kSortTile
// consecutive pairs (kSortItems per lane, read wave-striped: item j of lane
// l of wave w is element w * 64 * kSortItems + j * 64 + l of the tile, so
// each load instruction reads 64 consecutive keys). The pass kernel ranks a
// tile's pairs stably in that order (per wave: match masks from 8 ballots
// and a running per-digit count in LDS, its adds issued back to back; then
// the waves' counts are offset per digit), stages the tile in LDS in sorted
// order and writes it out from there, so consecutive lanes write
// consecutive addresses of each digit's run. A tile's output position per
// digit comes
// - for sorts of at most kChainedTiles tiles, from one digit count of all
//   passes over the input (k_radix_digits) and a decoupled look-back over
//   the tiles before it (one lane per digit, kLookBack predecessors per
//   round; status words tagged with the call's epoch as in the scan): one
//   launch per pass;
// - for larger sorts, from the tiles' digit counts of the pass
//   (k_radix_count, digit-major) and their exclusive scan: three launches
//   per pass, none of which waits on another workgroup (a look-back chain
//   through hundreds of tiles that start together costs more than the
//   extra read of the keys).
// Stable: within a tile the rank follows the element order, and tiles take
// their digit offsets in tile order. A chained tile only waits on
// lower-numbered tiles, which publish their counts before they wait (see
// the scan).

constexpr int kSortThreads = 512;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems; // 4096 pairs per tile
constexpr int kMaxPasses = 8;
constexpr int kDigitBlocks = 2048; // workgroups of the digit count, at most
constexpr int kDigitCopies = 16;   // the digit counts are spread over 16 copies (workgroup b adds to copy b % 16)
constexpr int64_t kChainedTiles = 32; // sorts of at most this many tiles take the chained passes

__device__ __forceinline__ uint32_t digit_of(uint64_t k, int shift) { return (uint32_t)(k >> shift) & (kRadix - 1); }

__global__ void k_radix_zero(uint32_t* p, int n)
{
	int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n)
		p[i] = 0u;
}

// Lanes of this wave that hold digit d (among the valid lanes).
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid)
{
	uint64_t same = __ballot(valid);
#pragma unroll
	for (int b = 0; b < kRadixBits; ++b)
	{
		bool bit = (d >> b) & 1u;
		uint64_t m = __ballot(bit);
		same &= bit ? m : ~m;
	}
	return same;
}

// Keys whose digit p is d, summed over the copies c of
// hist[(c * passes + p) * 256 + d]. Each workgroup counts chunks of 1024 keys
// (4 per lane) in LDS; a wave whose keys share one digit (a pass over bits
// the keys have in common) adds once, other waves add per lane. Workgroups
// add their counts to one of kDigitCopies copies: global adds to one address
// serialize, so one copy would take a few thousand in a row.
__global__ void __launch_bounds__(256) k_radix_digits(const uint64_t* keys, int64_t n, int passes, uint32_t* hist)
{
	__shared__ uint32_t cnt[kMaxPasses * kRadix];
	for (int k = threadIdx.x; k < passes * kRadix; k += 256)
		cnt[k] = 0u;
	__syncthreads();
	const int lane = threadIdx.x & 63;
	const int64_t chunks = (n + 1023) / 1024;
	for (int64_t c = blockIdx.x; c < chunks; c += gridDim.x)
	{
		uint64_t key[4];
		bool valid[4];
#pragma unroll
		for (int q = 0; q < 4; ++q)
		{
			int64_t i = c * 1024 + q * 256 + threadIdx.x;
			valid[q] = i < n;
			key[q] = valid[q] ? keys[i] : 0ull;
		}
#pragma unroll
		for (int q = 0; q < 4; ++q)
		{
			uint64_t live = __ballot(valid[q]);
			if (live == 0ull)
				continue;
			const int first = __ffsll((unsigned long long)live) - 1;
			for (int p = 0; p < passes; ++p)
			{
				uint32_t d = digit_of(key[q], p * kRadixBits);
				uint32_t d0 = (uint32_t)__shfl((int)d, first);
				if (__ballot(valid[q] && d != d0) == 0ull)
				{
					if (lane == first)
						atomicAdd(&cnt[p * kRadix + d0], (uint32_t)__popcll(live));
				}
				else if (valid[q])
					atomicAdd(&cnt[p * kRadix + d], 1u);
			}
		}
	}
	__syncthreads();
	uint32_t* copy = hist + (size_t)(blockIdx.x % kDigitCopies) * passes * kRadix;
	for (int k = threadIdx.x; k < passes * kRadix; k += 256)
		if (cnt[k] != 0u)
			atomicAdd(&copy[k], cnt[k]);
}

// Exclusive scans of two values over the first 256 lanes of the workgroup
// (lane t holds digit t's values; the other lanes pass zeros); tmp holds 8
// words. Every lane of the workgroup calls it.
__device__ __forceinline__ void block_scan2(uint32_t a, uint32_t b, uint32_t* tmp, uint32_t& ea, uint32_t& eb)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	uint32_t ia = wave_inclusive_sum(a), ib = wave_inclusive_sum(b);
	if (lane == 63 && wave < 4)
	{
		tmp[wave] = ia;
		tmp[4 + wave] = ib;
	}
	__syncthreads();
	uint32_t oa = 0, ob = 0;
	for (int w = 0; w < wave && w < 4; ++w)
	{
		oa += tmp[w];
		ob += tmp[4 + w];
	}
	ea = oa + ia - a;
	eb = ob + ib - b;
}

constexpr int kLookBack = 16; // predecessor tiles read per look-back round

// Per-tile digit counts of one pass, digit-major: counts[d * tiles + tile]
// (their exclusive scan is each tile's output position per digit).
__global__ void __launch_bounds__(kSortThreads) k_radix_count(const uint64_t* keys, int64_t n, int shift,
															 uint32_t* counts)
{
	__shared__ uint32_t cnt[kRadix];
	const int t = threadIdx.x;
	const int lane = t & 63, wave = t >> 6;
	const int64_t tile = blockIdx.x;
	for (int k = t; k < kRadix; k += kSortThreads)
		cnt[k] = 0u;
	uint64_t key[kSortItems];
	const int64_t waveBase = tile * kSortTile + (int64_t)wave * 64 * kSortItems;
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		int64_t i = waveBase + j * 64 + lane;
		key[j] = i < n ? keys[i] : 0ull;
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		bool valid = waveBase + j * 64 + lane < n;
		uint64_t live = __ballot(valid);
		if (live == 0ull)
			continue;
		const int first = __ffsll((unsigned long long)live) - 1;
		uint32_t d = digit_of(key[j], shift);
		uint32_t d0 = (uint32_t)__shfl((int)d, first);
		if (__ballot(valid && d != d0) == 0ull)
		{
			if (lane == first)
				atomicAdd(&cnt[d0], (uint32_t)__popcll(live));
		}
		else if (valid)
			atomicAdd(&cnt[d], 1u);
	}
	__syncthreads();
	for (int k = t; k < kRadix; k += kSortThreads)
		counts[(int64_t)k * gridDim.x + tile] = cnt[k];
}

// Chained: digitCount is the digit's count over the input in kDigitCopies
// copies copyStride apart, and the tile finds its offsets by look-back.
// Otherwise digitCount holds the scanned per-tile counts (k_radix_count):
// the tile's output position of each digit.
template <bool Chained>
__global__ void __launch_bounds__(kSortThreads) k_radix_pass(const uint64_t* keys, const uint32_t* vals,
															uint64_t* keysOut, uint32_t* valsOut, int64_t n, int shift,
															const uint32_t* digitCount, int copyStride,
															uint64_t* status, uint32_t epoch)
{
	__shared__ uint64_t sKey[kSortTile];
	__shared__ uint32_t sVal[kSortTile];
	__shared__ uint32_t waveCnt[kSortWaves][kRadix];
	__shared__ uint32_t scanTmp[8];
	__shared__ uint32_t tileStart[kRadix]; // tile-local start of each digit
	__shared__ int64_t outBase[kRadix];    // global position of the tile's first key of each digit, minus tileStart
	const int t = threadIdx.x;
	const int lane = t & 63, wave = t >> 6;
	const int64_t tile = blockIdx.x;
	const int64_t tileBase = tile * kSortTile;
	for (int k = t; k < kSortWaves * kRadix; k += kSortThreads)
		(&waveCnt[0][0])[k] = 0u;

	// the digit's count over the whole input (lane t < 256), read with the keys
	uint32_t copies[Chained ? kDigitCopies : 1];
	if constexpr (Chained)
	{
#pragma unroll
		for (int c = 0; c < kDigitCopies; ++c)
			copies[c] = t < kRadix ? digitCount[c * copyStride + t] : 0u;
	}
	else
		copies[0] = t < kRadix ? digitCount[(int64_t)t * gridDim.x + tile] : 0u;
	uint64_t key[kSortItems];
	uint32_t val[kSortItems];
	const int64_t waveBase = tileBase + (int64_t)wave * 64 * kSortItems;
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		int64_t i = waveBase + j * 64 + lane;
		bool valid = i < n;
		key[j] = valid ? keys[i] : 0ull;
		val[j] = valid ? vals[i] : 0u;
	}
	__syncthreads();
	// Stable rank within the wave, in item order (j, then lane): the
	// lowest lane of each digit group adds the group's size to the wave's
	// count of that digit and gets the count before it. The adds of
	// successive items are issued back to back; LDS executes one wave's
	// operations in order, so each returns the count its item follows.
	const uint64_t lowerMask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	uint32_t below[kSortItems], leader[kSortItems], old[kSortItems];
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		bool valid = waveBase + j * 64 + lane < n;
		uint32_t d = digit_of(key[j], shift);
		uint64_t same = digit_peers(d, valid);
		below[j] = (uint32_t)__popcll(same & lowerMask);
		leader[j] = valid ? (uint32_t)(__ffsll((unsigned long long)same) - 1) : (uint32_t)lane;
		old[j] = 0u;
		if (valid && below[j] == 0u)
			old[j] = atomicAdd(&waveCnt[wave][d], (uint32_t)__popcll(same));
	}
	uint32_t rank[kSortItems];
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
		rank[j] = (uint32_t)__shfl((int)old[j], (int)leader[j]) + below[j];
	__syncthreads();
	// lane t < 256 owns digit t: the waves' offsets, the tile's count
	uint32_t tileCount = 0;
	if (t < kRadix)
	{
		uint32_t c[kSortWaves];
#pragma unroll
		for (int w = 0; w < kSortWaves; ++w)
			c[w] = waveCnt[w][t];
#pragma unroll
		for (int w = 0; w < kSortWaves; ++w)
		{
			waveCnt[w][t] = tileCount;
			tileCount += c[w];
		}
	}
	// tile-local start of the digit; keys of smaller digits in the whole input
	uint32_t total = 0;
	if constexpr (Chained)
	{
#pragma unroll
		for (int c = 0; c < kDigitCopies; ++c)
			total += copies[c];
	}
	uint32_t start, allBefore;
	block_scan2(tileCount, total, scanTmp, start, allBefore);
	if (t < kRadix)
		tileStart[t] = start;
	__syncthreads();
	// stage in tile order of the digit (while the lanes of the digits look
	// back), then write out
#pragma unroll
	for (int j = 0; j < kSortItems; ++j)
	{
		if (waveBase + j * 64 + lane < n)
		{
			uint32_t d = digit_of(key[j], shift);
			uint32_t at = tileStart[d] + waveCnt[wave][d] + rank[j];
			sKey[at] = key[j];
			sVal[at] = val[j];
		}
	}
	if (!Chained && t < kRadix)
		outBase[t] = (int64_t)copies[0] - (int64_t)start;
	if (Chained && t < kRadix)
	{
		// this digit's keys in the tiles before this one: decoupled
		// look-back, kLookBack predecessors per round (their aggregates are
		// summed up to the nearest inclusive prefix; a predecessor not yet
		// published ends the round early and is read again)
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
			int64_t j = tile - 1;
			for (;;)
			{
				uint64_t word[kLookBack];
#pragma unroll
				for (int q = 0; q < kLookBack; ++q)
					word[q] = j - q >= 0 ? __hip_atomic_load(&status[(j - q) * kRadix + t], __ATOMIC_RELAXED,
															 __HIP_MEMORY_SCOPE_AGENT)
										 : pack_status(epoch, kFlagPrefix, 0u);
				int used = 0;
				bool done = false;
#pragma unroll
				for (int q = 0; q < kLookBack; ++q)
				{
					if (done || used < q)
						break;
					uint64_t w = word[q];
					uint64_t flag = (w >> 32) & 3ull;
					if ((uint32_t)(w >> 34) != epoch || flag == 0ull)
						break;
					prefix += (uint32_t)w;
					used = q + 1;
					done = flag == kFlagPrefix;
				}
				if (done)
					break;
				j -= used;
				if (used < kLookBack)
					__builtin_amdgcn_s_sleep(1);
			}
			__hip_atomic_store(&st[t], pack_status(epoch, kFlagPrefix, prefix + tileCount), __ATOMIC_RELAXED,
							   __HIP_MEMORY_SCOPE_AGENT);
		}
		outBase[t] = (int64_t)allBefore + (int64_t)prefix - (int64_t)start;
	}
	__syncthreads();
	const int tileN = (int)(n - tileBase < kSortTile ? n - tileBase : kSortTile);
#pragma unroll
	for (int q = 0; q < kSortItems; ++q)
	{
		int at = q * kSortThreads + t;
		if (at < tileN)
		{
			uint64_t k = sKey[at];
			int64_t dst = outBase[digit_of(k, shift)] + at;
			keysOut[dst] = k;
			valsOut[dst] = sVal[at];
		}
	}
}

size_t sort_scratch_bytes(int64_t n)
{
	int64_t tiles = (n + kSortTile - 1) / kSortTile;
	return std::max((size_t)kDigitCopies * kMaxPasses * kRadix, (size_t)tiles * kRadix) * sizeof(uint32_t);
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
	if (radix_sort_pairs_in_place_or_alt(keys, vals, keysAlt, valsAlt, n, bits, scratch, status, s))
	{
		hipMemcpyAsync(keys, keysAlt, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
		hipMemcpyAsync(vals, valsAlt, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
	}
}

bool radix_sort_pairs_in_place_or_alt(uint64_t* keys, uint32_t* vals, uint64_t* keysAlt, uint32_t* valsAlt, int64_t n,
									  int bits, void* scratch, uint64_t* status, hipStream_t s)
{
	if (n <= 1 || bits <= 0)
	{
		return false;
	}
	const int passes = std::min((bits + kRadixBits - 1) / kRadixBits, kMaxPasses);
	const int64_t tiles = (n + kSortTile - 1) / kSortTile;
	uint32_t* hist = (uint32_t*)scratch;
	// algorithmic bytes: the keys read once for the digit counts (8 B), then
	// every (key, value) pair read and written once per pass (24 B; the
	// per-tile counts of larger sorts read the keys once more per pass)
	const SortTimer* timer = t_sortTimer;
	if (timer != nullptr)
		timer->begin(timer->ctx, 8.0 * (double)n + 24.0 * (double)n * passes);
	uint64_t* kin = keys;
	uint32_t* vin = vals;
	uint64_t* kout = keysAlt;
	uint32_t* vout = valsAlt;
	if (tiles <= kChainedTiles)
	{
		// few tiles: one launch per pass, offsets by look-back from the digit
		// counts of all passes
		const int histWords = kDigitCopies * passes * kRadix;
		k_radix_zero<<<(histWords + 255) / 256, 256, 0, s>>>(hist, histWords);
		const int64_t histBlocks = std::min<int64_t>((n + 1023) / 1024, kDigitBlocks);
		k_radix_digits<<<(unsigned)histBlocks, 256, 0, s>>>(keys, n, passes, hist);
		for (int p = 0; p < passes; ++p)
		{
			k_radix_pass<true><<<(unsigned)tiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, p * kRadixBits,
																		hist + p * kRadix, passes * kRadix, status,
																		next_scan_epoch());
			std::swap(kin, kout);
			std::swap(vin, vout);
		}
	}
	else
	{
		// many tiles (a look-back chain through tiles that run at once would
		// set the pace): per pass, the tiles' digit counts, their scan, and
		// the pass from the scanned offsets
		uint32_t* offs = (uint32_t*)scratch;
		for (int p = 0; p < passes; ++p)
		{
			k_radix_count<<<(unsigned)tiles, kSortThreads, 0, s>>>(kin, n, p * kRadixBits, offs);
			scan_exclusive(offs, offs, tiles * kRadix, nullptr, status, s);
			k_radix_pass<false><<<(unsigned)tiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, p * kRadixBits,
																		 offs, 0, status, 0u);
			std::swap(kin, kout);
			std::swap(vin, vout);
		}
	}
	if (timer != nullptr)
		timer->end(timer->ctx);
	return kin != keys;
}

} // namespace mx
