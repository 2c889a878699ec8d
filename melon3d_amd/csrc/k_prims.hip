// Scan and stable radix sort for gfx950: 256-thread workgroups (4 waves of
// 64 lanes). The sort ranks each element inside its workgroup with wave
// ballots (one ballot per digit bit) plus a per-wave digit table in LDS, so
// the scatter is stable without per-element atomics.
#include "mx_prims.h"

#include <stdio.h>

#include <atomic>

namespace mx
{

constexpr int kScanBlock = 1024; // elements per scan workgroup (4 per thread)
constexpr int kSortBlock = 256;  // elements per sort workgroup (1 per thread)
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

__device__ __forceinline__ uint32_t digit_of(uint64_t k, int shift) { return (uint32_t)(k >> shift) & (kRadix - 1); }

// hist layout: [digit][block] so one linear exclusive scan yields the
// global scatter base of every (digit, block).
__global__ void __launch_bounds__(256) k_radix_hist(const uint64_t* keys, int64_t n, int shift, uint32_t* hist,
													int nb)
{
	__shared__ uint32_t cnt[kRadix];
	cnt[threadIdx.x] = 0;
	__syncthreads();
	int64_t i = (int64_t)blockIdx.x * kSortBlock + threadIdx.x;
	if (i < n)
	{
		atomicAdd(&cnt[digit_of(keys[i], shift)], 1u);
	}
	__syncthreads();
	hist[(int64_t)threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_radix_scatter(const uint64_t* keys, const uint32_t* vals, uint64_t* keysOut,
													   uint32_t* valsOut, int64_t n, int shift,
													   const uint32_t* base, int nb)
{
	__shared__ uint32_t waveCnt[4][kRadix];
	const int t = threadIdx.x;
	const int wave = t >> 6;
	const int lane = t & 63;
	for (int k = t; k < 4 * kRadix; k += 256)
	{
		(&waveCnt[0][0])[k] = 0;
	}
	__syncthreads();

	int64_t i = (int64_t)blockIdx.x * kSortBlock + t;
	bool valid = i < n;
	uint64_t key = valid ? keys[i] : 0;
	uint32_t val = valid ? vals[i] : 0;
	uint32_t d = digit_of(key, shift);

	// lanes of this wave holding the same digit
	uint64_t same = __ballot(valid);
	for (int b = 0; b < kRadixBits; ++b)
	{
		bool bit = (d >> b) & 1u;
		uint64_t m = __ballot(bit);
		same &= bit ? m : ~m;
	}
	uint64_t lowerMask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	uint32_t rankInWave = (uint32_t)__popcll(same & lowerMask);
	if (valid && rankInWave == 0)
	{
		waveCnt[wave][d] = (uint32_t)__popcll(same);
	}
	__syncthreads();
	if (valid)
	{
		uint32_t r = rankInWave;
		for (int w = 0; w < wave; ++w)
		{
			r += waveCnt[w][d];
		}
		uint32_t dst = base[(int64_t)d * nb + blockIdx.x] + r;
		keysOut[dst] = key;
		valsOut[dst] = val;
	}
}

size_t sort_scratch_bytes(int64_t n)
{
	int64_t nb = (n + kSortBlock - 1) / kSortBlock;
	int64_t h = nb * kRadix;
	return (size_t)h * sizeof(uint32_t);
}

int64_t sort_status_words(int64_t n)
{
	int64_t nb = (n + kSortBlock - 1) / kSortBlock;
	return scan_status_words(nb * kRadix);
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
	int nb = (int)((n + kSortBlock - 1) / kSortBlock);
	int64_t h = (int64_t)nb * kRadix;
	uint32_t* hist = (uint32_t*)scratch;
	int passes = (bits + kRadixBits - 1) / kRadixBits;
	// algorithmic bytes per pass: the keys read for the histogram (8 B), then
	// every (key, value) pair read and written once (24 B)
	const SortTimer* timer = t_sortTimer;
	if (timer != nullptr)
		timer->begin(timer->ctx, 32.0 * (double)n * passes);
	uint64_t* kin = keys;
	uint32_t* vin = vals;
	uint64_t* kout = keysAlt;
	uint32_t* vout = valsAlt;
	for (int p = 0; p < passes; ++p)
	{
		int shift = p * kRadixBits;
		k_radix_hist<<<nb, 256, 0, s>>>(kin, n, shift, hist, nb);
		scan_exclusive(hist, hist, h, nullptr, status, s);
		k_radix_scatter<<<nb, 256, 0, s>>>(kin, vin, kout, vout, n, shift, hist, nb);
		uint64_t* tk = kin;
		kin = kout;
		kout = tk;
		uint32_t* tv = vin;
		vin = vout;
		vout = tv;
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
