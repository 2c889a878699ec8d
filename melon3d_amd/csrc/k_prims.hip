// Scan and stable radix sort for gfx950: 256-thread workgroups (4 waves of
// 64 lanes). The sort ranks each element inside its workgroup with wave
// ballots (one ballot per digit bit) plus a per-wave digit table in LDS, so
// the scatter is stable without per-element atomics.
#include "mx_prims.h"

#include <stdio.h>

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

__global__ void __launch_bounds__(256) k_scan_reduce(const uint32_t* in, int64_t n, uint32_t* blockSums)
{
	__shared__ uint32_t lds[256];
	int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * 4;
	uint32_t s = 0;
	for (int k = 0; k < 4; ++k)
	{
		if (base + k < n)
		{
			s += in[base + k];
		}
	}
	uint32_t tot;
	block_exclusive_sum(s, lds, &tot);
	if (threadIdx.x == 0)
	{
		blockSums[blockIdx.x] = tot;
	}
}

// Scans up to 256*1024 block sums in one workgroup (serially in chunks).
__global__ void __launch_bounds__(256) k_scan_blocksums(uint32_t* sums, int64_t nb, uint32_t* total)
{
	__shared__ uint32_t lds[256];
	uint32_t carry = 0;
	for (int64_t base = 0; base < nb; base += 256)
	{
		int64_t i = base + threadIdx.x;
		uint32_t v = i < nb ? sums[i] : 0u;
		uint32_t tot;
		uint32_t ex = block_exclusive_sum(v, lds, &tot);
		if (i < nb)
		{
			sums[i] = ex + carry;
		}
		carry += tot;
	}
	if (threadIdx.x == 0 && total != nullptr)
	{
		*total = carry;
	}
}

__global__ void __launch_bounds__(256) k_scan_final(const uint32_t* in, uint32_t* out, int64_t n,
													const uint32_t* blockOffsets)
{
	__shared__ uint32_t lds[256];
	int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * 4;
	uint32_t v[4];
	uint32_t s = 0;
	for (int k = 0; k < 4; ++k)
	{
		v[k] = base + k < n ? in[base + k] : 0u;
		s += v[k];
	}
	uint32_t tot;
	uint32_t ex = block_exclusive_sum(s, lds, &tot) + blockOffsets[blockIdx.x];
	for (int k = 0; k < 4; ++k)
	{
		if (base + k < n)
		{
			out[base + k] = ex;
		}
		ex += v[k];
	}
}

size_t scan_scratch_bytes(int64_t n)
{
	int64_t nb = (n + kScanBlock - 1) / kScanBlock;
	return (size_t)(nb + 1) * sizeof(uint32_t) + 256;
}

void scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, void* scratch, hipStream_t s)
{
	if (n <= 0)
	{
		if (total != nullptr)
		{
			hipMemsetAsync(total, 0, sizeof(uint32_t), s);
		}
		return;
	}
	int64_t nb = (n + kScanBlock - 1) / kScanBlock;
	uint32_t* sums = (uint32_t*)scratch;
	k_scan_reduce<<<(unsigned)nb, 256, 0, s>>>(in, n, sums);
	k_scan_blocksums<<<1, 256, 0, s>>>(sums, nb, total);
	k_scan_final<<<(unsigned)nb, 256, 0, s>>>(in, out, n, sums);
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
	return (size_t)h * sizeof(uint32_t) + scan_scratch_bytes(h) + 256;
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keysAlt, uint32_t* valsAlt, int64_t n, int bits,
					  void* scratch, hipStream_t s)
{
	if (n <= 1 || bits <= 0)
	{
		return;
	}
	int nb = (int)((n + kSortBlock - 1) / kSortBlock);
	int64_t h = (int64_t)nb * kRadix;
	uint32_t* hist = (uint32_t*)scratch;
	void* scanScratch = (void*)((char*)scratch + ((h * sizeof(uint32_t) + 255) & ~(size_t)255));
	int passes = (bits + kRadixBits - 1) / kRadixBits;
	uint64_t* kin = keys;
	uint32_t* vin = vals;
	uint64_t* kout = keysAlt;
	uint32_t* vout = valsAlt;
	for (int p = 0; p < passes; ++p)
	{
		int shift = p * kRadixBits;
		k_radix_hist<<<nb, 256, 0, s>>>(kin, n, shift, hist, nb);
		scan_exclusive(hist, hist, h, nullptr, scanScratch, s);
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
}

} // namespace mx
