// Broadphase and contact-set maintenance kernels.
//
//   refresh     fat-AABB breach test and tier flags, one thread per body
//               (src/world.cpp:156-229)
//   tiers       count -> scan -> emit of (cell key, body) entries in body
//               order, then a stable radix sort on a compressed cell key,
//               which yields the reference's unique (key, body) order
//               (src/grid.cpp:39-131)
//   pairs       one thread per cell run of the active tier enumerates the
//               run's candidates in the reference's serial order (active x
//               active, then active x stable), count pass + emit pass;
//               coarse-list candidates follow in list order
//               (src/world.cpp:1570-1702)
//   create      first admitted occurrence of each pair key wins, preserving
//               candidate order (the serial creation re-check)
//   remove      separated awake pairs are swap-removed in descending index
//               order; the k-th largest removed slot below the new size
//               receives the k-th largest surviving slot above it
//               (src/world.cpp:1705-1735, :100-119)
#include "mx_steps.h"
#include "mx_wave.h"

#include <algorithm>
#include <string.h>

namespace mx
{

constexpr int kBlock = 256;

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, int64_t n)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s)
{
	if (n > 0)
		k_fill_i32<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
}
void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
	if (n > 0)
		k_fill_u32<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
}
static void launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s)
{
	if (n > 0)
		k_fill_u64<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
}

// Up to four fills of 32-bit words in one launch (a fill is a launch of a
// few microseconds; several small ones in a row cost more than the work).
__global__ void k_fill_multi(FillSegs F)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	for (int k = 0; k < F.count; ++k)
	{
		if (i < F.n[k])
			F.p[k][i] = F.v[k];
	}
}

void launch_fill_multi(const FillSegs& F, hipStream_t s)
{
	int64_t m = 0;
	for (int k = 0; k < F.count; ++k)
		m = std::max(m, F.n[k]);
	if (m > 0)
		k_fill_multi<<<grid_for(m), kBlock, 0, s>>>(F);
}

static DBodies body_view(World& w)
{
	DBodies B;
	B.pos = w.bc.pos.p;
	B.rot = w.bc.rot.p;
	B.vel = w.bc.vel.p;
	B.angVel = w.bc.angVel.p;
	B.prevPos = w.bc.prevPos.p;
	B.prevRot = w.bc.prevRot.p;
	B.prevVel = w.bc.prevVel.p;
	B.prevAngVel = w.bc.prevAngVel.p;
	B.invIL = w.bc.invIL.p;
	B.iw0 = w.bc.iw0.p;
	B.iw1 = w.bc.iw1.p;
	B.iw2 = w.bc.iw2.p;
	B.shape = w.bc.shape.p;
	B.mat = w.bc.mat.p;
	B.force = w.bc.force.p;
	B.torque = w.bc.torque.p;
	B.flags = w.bc.flags.p;
	B.filter = w.bc.filter.p;
	B.fatLo = w.bc.fatLo.p;
	B.fatHi = w.bc.fatHi.p;
	B.bp = w.bc.bp.p;
	B.sb = w.bc.sb.p;
	B.sub = w.subCount > 0 ? w.bodySub.p : nullptr;
	B.subStable = w.subCount > 0 ? w.subStable.p : nullptr;
	return B;
}

DContacts columns_view(ContactColumns& cc, int32_t cap)
{
	DContacts C;
	C.cap = cap;
	C.cs = 4;
	C.js = 1;
	C.bodies = cc.bodies.p;
	C.mat = cc.mat.p;
	C.normal = cc.normal.p;
	C.cache = cc.cache.p;
	C.pA = cc.pA.p;
	C.pB = cc.pB.p;
	C.pJA = cc.pJA.p;
	C.pJB = cc.pJB.p;
	C.pRA = cc.pRA.p;
	C.pRB = cc.pRB.p;
	C.pFA = cc.pFA.p;
	C.pFB = cc.pFB.p;
	C.pSlip = cc.pSlip.p;
	C.hint = cc.hint.p;
	return C;
}

static DContacts contact_view(World& w) { return columns_view(w.cc, w.contactCap); }

DBodies make_body_view(World& w) { return body_view(w); }
DContacts make_contact_view(World& w) { return contact_view(w); }

// ------------------------------------------------------------ refresh

MX_HD m3d_aabb fat_lo_hi(float4 lo, float4 hi)
{
	m3d_aabb a;
	a.lowerBound = f4_v3(lo);
	a.upperBound = f4_v3(hi);
	return a;
}

// One body's broadphase refresh; true when its fat AABB was refitted.
__device__ __forceinline__ bool refresh_body(const DBodies& B, int32_t i, float dt, int rebuildStable)
{
	bool refit = false;
	uint32_t fl = B.flags[i];
	bool inUse = (fl & BF_IN_USE) != 0;
	bool awake = (fl & BF_AWAKE) != 0;
	int type = bf_type(fl);
	uint8_t bp = B.bp[i] & (BP_HAS_FAT | BP_IN_STABLE_TIER | BP_IN_STABLE_CELLS);
	bool stable = inUse && !awake;
	bool active = inUse && (awake || !(bp & BP_IN_STABLE_TIER));
	if (active)
		bp |= BP_ACTIVE;
	if (stable)
		bp |= BP_STABLE;
	if (B.sub != nullptr)
		rebuildStable = (B.subStable[B.sub[i]] & SUB_REBUILD) != 0;
	uint8_t sb = (uint8_t)((inUse ? SB_IN_USE : 0) | ((inUse && type == M3D_BODY_DYNAMIC) ? SB_DYNAMIC : 0) |
						   ((inUse && type == M3D_BODY_KINEMATIC) ? SB_KINEMATIC : 0) | (awake ? SB_AWAKE : 0));
	if (active)
	{
		Shape sh = load_shape(B, i);
		m3d_transform xf = load_xf(B, i);
		m3d_vec3 v = f4_v3(B.vel[i]);
		m3d_aabb pred = shape_aabb(sh, xf);
		m3d_vec3 st = m3d_scale(dt, v);
		if (st.x < 0.0f) pred.lowerBound.x += st.x; else pred.upperBound.x += st.x;
		if (st.y < 0.0f) pred.lowerBound.y += st.y; else pred.upperBound.y += st.y;
		if (st.z < 0.0f) pred.lowerBound.z += st.z; else pred.upperBound.z += st.z;
		if (!(bp & BP_HAS_FAT) || !m3d_aabb_contains(fat_lo_hi(B.fatLo[i], B.fatHi[i]), pred))
		{
			m3d_vec3 mg = m3d_v3(kAABBMargin, kAABBMargin, kAABBMargin);
			m3d_aabb fat;
			fat.lowerBound = m3d_sub(pred.lowerBound, mg);
			fat.upperBound = m3d_add(pred.upperBound, mg);
			m3d_vec3 sl = m3d_scale(2.0f * dt, v);
			if (sl.x < 0.0f) fat.lowerBound.x += sl.x; else fat.upperBound.x += sl.x;
			if (sl.y < 0.0f) fat.lowerBound.y += sl.y; else fat.upperBound.y += sl.y;
			if (sl.z < 0.0f) fat.lowerBound.z += sl.z; else fat.upperBound.z += sl.z;
			B.fatLo[i] = v3_f4(fat.lowerBound, 0.0f);
			B.fatHi[i] = v3_f4(fat.upperBound, 0.0f);
			bp |= BP_HAS_FAT | BP_MOVED;
			refit = true;
		}
	}
	else if (stable && rebuildStable && !(bp & BP_HAS_FAT))
	{
		// ExpandedAABB: margin first, then the one-step velocity extension
		Shape sh = load_shape(B, i);
		m3d_transform xf = load_xf(B, i);
		m3d_aabb a = shape_aabb(sh, xf);
		m3d_vec3 mg = m3d_v3(kAABBMargin, kAABBMargin, kAABBMargin);
		a.lowerBound = m3d_sub(a.lowerBound, mg);
		a.upperBound = m3d_add(a.upperBound, mg);
		m3d_vec3 d = m3d_scale(dt, f4_v3(B.vel[i]));
		if (d.x < 0.0f) a.lowerBound.x += d.x; else a.upperBound.x += d.x;
		if (d.y < 0.0f) a.lowerBound.y += d.y; else a.upperBound.y += d.y;
		if (d.z < 0.0f) a.lowerBound.z += d.z; else a.upperBound.z += d.z;
		B.fatLo[i] = v3_f4(a.lowerBound, 0.0f);
		B.fatHi[i] = v3_f4(a.upperBound, 0.0f);
		bp |= BP_HAS_FAT;
	}
	B.bp[i] = bp;
	B.sb[i] = sb;
	return refit;
}

// The refit count is summed per workgroup (one atomic each) over a grid of
// at most kRefreshGroups workgroups: an atomic from every wave (or from
// every workgroup of a 200k-body grid) on one counter serializes.
constexpr unsigned kRefreshGroups = 256;
__global__ void k_refresh(DBodies B, int32_t n, float dt, int rebuildStable, uint32_t* counters)
{
	uint32_t refits = 0;
	for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int32_t)(gridDim.x * blockDim.x))
		refits += refresh_body(B, i, dt, rebuildStable) ? 1u : 0u;
	block_add<1>(counters, { CT_REFAT }, { refits });
}

void launch_refresh(World& w, float dt, bool rebuildStable)
{
	int32_t n = w.deviceBodyCount;
	if (n > 0)
	{
		const unsigned g = std::min<unsigned>(grid_for(n), kRefreshGroups);
		k_refresh<<<g, kBlock, 0, w.stream>>>(body_view(w), n, dt, rebuildStable ? 1 : 0, w.counters.p);
	}
}

// inStableTier := isStable (src/world.cpp:215-218), for the sub-worlds that
// rebuild their stable tier when worlds are batched
__global__ void k_stable_migrate(DBodies B, int32_t n)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	if (B.sub != nullptr && !(B.subStable[B.sub[i]] & SUB_REBUILD))
		return;
	uint8_t b = B.bp[i];
	B.bp[i] = (b & BP_STABLE) ? (uint8_t)(b | BP_IN_STABLE_TIER) : (uint8_t)(b & ~BP_IN_STABLE_TIER);
}

void launch_stable_migrate(World& w)
{
	int32_t n = w.deviceBodyCount;
	if (n > 0)
		k_stable_migrate<<<grid_for(n), kBlock, 0, w.stream>>>(body_view(w), n);
}

// -------------------------------------------------------------- tiers


MX_HD uint64_t cell_key_of(m3d_vec3 p, float cellSize)
{
	float inv = 1.0f / cellSize;
	return pack_cell(cell_coord(p.x, inv), cell_coord(p.y, inv), cell_coord(p.z, inv));
}

// Batched worlds: the sub-world index above three 15-bit cell fields, so
// cells of different sub-worlds never share a key and each sub-world's cells
// keep the reference's x, y, z order (exact within 2^14 cells of the origin,
// where the reference's 21-bit fields are exact too).
MX_HD uint64_t pack_cell_sub(uint32_t s, int64_t x, int64_t y, int64_t z)
{
	const int64_t off = (int64_t)1 << 14;
	uint64_t ux = (uint64_t)(x + off) & 0x7FFFu;
	uint64_t uy = (uint64_t)(y + off) & 0x7FFFu;
	uint64_t uz = (uint64_t)(z + off) & 0x7FFFu;
	return ((uint64_t)s << 45) | (ux << 30) | (uy << 15) | uz;
}

__device__ __forceinline__ uint64_t body_cell_key(const DBodies& B, int32_t body, int64_t x, int64_t y, int64_t z)
{
	return B.sub != nullptr ? pack_cell_sub((uint32_t)B.sub[body], x, y, z) : pack_cell(x, y, z);
}

// cell fields of a stored key: sub-world, x, y, z
struct KeyFields
{
	uint32_t s, x, y, z;
};

MX_HD KeyFields key_fields(uint64_t k, int batched)
{
	KeyFields f;
	if (batched)
	{
		f.s = (uint32_t)(k >> 45);
		f.x = (uint32_t)(k >> 30) & 0x7FFFu;
		f.y = (uint32_t)(k >> 15) & 0x7FFFu;
		f.z = (uint32_t)k & 0x7FFFu;
	}
	else
	{
		f.s = 0;
		f.x = (uint32_t)(k >> 42) & 0x1FFFFFu;
		f.y = (uint32_t)(k >> 21) & 0x1FFFFFu;
		f.z = (uint32_t)k & 0x1FFFFFu;
	}
	return f;
}

// body included in a tier build: its tier bit, and (batched stable builds)
// its sub-world rebuilds this step
__device__ __forceinline__ bool tier_included(const DBodies& B, int32_t i, uint8_t includeBit, int gated)
{
	return (B.bp[i] & includeBit) && (!gated || (B.subStable[B.sub[i]] & SUB_REBUILD));
}

__global__ void k_tier_count(DBodies B, int32_t n, uint8_t includeBit, int gated, float cellSize, uint32_t* cnt,
							 uint32_t* ov)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	uint32_t c = 0, o = 0;
	if (tier_included(B, i, includeBit, gated))
	{
		float inv = 1.0f / cellSize;
		float4 lo = B.fatLo[i], hi = B.fatHi[i];
		int64_t x0 = cell_coord(lo.x, inv), y0 = cell_coord(lo.y, inv), z0 = cell_coord(lo.z, inv);
		int64_t x1 = cell_coord(hi.x, inv), y1 = cell_coord(hi.y, inv), z1 = cell_coord(hi.z, inv);
		int64_t span = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
		if (span > kOversizedCellSpan)
			o = 1;
		else
			c = (uint32_t)span;
	}
	cnt[i] = c;
	ov[i] = o;
}

__global__ void k_compact_indices(const uint32_t* flags, const uint32_t* offs, int64_t n, int32_t* out)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flags[i])
		out[offs[i]] = (int32_t)i;
}

// *gmin = min(*gmin, mn) and *gmax = max(*gmax, mx) over the workgroup's
// lanes with valid set: wave shuffles, then one pair of global atomics per
// workgroup (callers loop over their elements on reduce_grid workgroups).
// With part set, the workgroup's (min, max) goes to part[0], part[1]
// instead (all ones and zero for none), for k_bounds_reduce: the atomics of
// hundreds of workgroups on one line serialize. Every lane of the workgroup
// calls it.
__device__ __forceinline__ void block_minmax(uint32_t mn, uint32_t mx, bool valid, uint32_t* gmin, uint32_t* gmax,
											 uint32_t* part = nullptr)
{
	__shared__ uint32_t smin[16], smax[16];
	mn = wave_min(valid ? mn : 0xFFFFFFFFu);
	mx = wave_max(valid ? mx : 0u);
	const uint32_t wv = threadIdx.x >> 6;
	if (__lane_id() == 0)
	{
		smin[wv] = mn;
		smax[wv] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		const uint32_t nw = (blockDim.x + 63) >> 6;
		for (uint32_t k = 1; k < nw; ++k)
		{
			mn = smin[k] < mn ? smin[k] : mn;
			mx = smax[k] > mx ? smax[k] : mx;
		}
		if (part != nullptr)
		{
			part[0] = mn;
			part[1] = mx;
		}
		else if (mn <= mx)
		{
			atomicMin(gmin, mn);
			atomicMax(gmax, mx);
		}
	}
	__syncthreads();
}

// Cell bounds from the per-workgroup partials of k_tier_emit: one
// workgroup, nb partials of six words (x min, x max, y min, y max, z min,
// z max), written to the counters block.
__global__ void __launch_bounds__(256) k_bounds_reduce(const uint32_t* part, int32_t nb, uint32_t* counters)
{
	__shared__ uint32_t red[6][4];
	uint32_t v[6] = { 0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u };
	for (int32_t b = (int32_t)threadIdx.x; b < nb; b += (int32_t)blockDim.x)
	{
#pragma unroll
		for (int k = 0; k < 6; ++k)
		{
			uint32_t x = part[6 * (int64_t)b + k];
			v[k] = (k & 1) ? (x > v[k] ? x : v[k]) : (x < v[k] ? x : v[k]);
		}
	}
	const uint32_t wv = threadIdx.x >> 6;
#pragma unroll
	for (int k = 0; k < 6; ++k)
	{
		v[k] = (k & 1) ? wave_max(v[k]) : wave_min(v[k]);
		if (__lane_id() == 0)
			red[k][wv] = v[k];
	}
	__syncthreads();
	if (threadIdx.x < 6u)
	{
		const int k = (int)threadIdx.x;
		const uint32_t nw = (blockDim.x + 63) >> 6;
		uint32_t r = red[k][0];
		for (uint32_t w = 1; w < nw; ++w)
			r = (k & 1) ? (red[k][w] > r ? red[k][w] : r) : (red[k][w] < r ? red[k][w] : r);
		counters[CT_UX_MIN + k] = r;
	}
}

// Entries at positions >= room are not stored (their cell bounds still
// count): the caller emits again into a larger buffer.
// part: per-workgroup bounds (six words each, then k_bounds_reduce), or null
// for atomics on the counters.
__global__ void k_tier_emit(DBodies B, int32_t n, uint8_t includeBit, int gated, float cellSize, const uint32_t* cnt,
							const uint32_t* offs, uint64_t* keys, uint32_t* bodies, int64_t room, uint32_t* counters,
							uint32_t* part)
{
	uint32_t mnx = 0xFFFFFFFFu, mxx = 0, mny = 0xFFFFFFFFu, mxy = 0, mnz = 0xFFFFFFFFu, mxz = 0;
	bool any = false;
	const int batched = B.sub != nullptr;
	for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int32_t)(gridDim.x * blockDim.x))
	{
		if (!(cnt[i] > 0 && tier_included(B, i, includeBit, gated)))
			continue;
		float inv = 1.0f / cellSize;
		float4 lo = B.fatLo[i], hi = B.fatHi[i];
		int64_t x0 = cell_coord(lo.x, inv), y0 = cell_coord(lo.y, inv), z0 = cell_coord(lo.z, inv);
		int64_t x1 = cell_coord(hi.x, inv), y1 = cell_coord(hi.y, inv), z1 = cell_coord(hi.z, inv);
		uint32_t o = offs[i];
		for (int64_t x = x0; x <= x1; ++x)
			for (int64_t y = y0; y <= y1; ++y)
				for (int64_t z = z0; z <= z1; ++z)
				{
					uint64_t k = body_cell_key(B, i, x, y, z);
					if ((int64_t)o < room)
					{
						keys[o] = k;
						bodies[o] = (uint32_t)i;
					}
					++o;
					KeyFields f = key_fields(k, batched);
					uint32_t ux = f.x, uy = f.y, uz = f.z;
					mnx = min(mnx, ux);
					mxx = max(mxx, ux);
					mny = min(mny, uy);
					mxy = max(mxy, uy);
					mnz = min(mnz, uz);
					mxz = max(mxz, uz);
					any = true;
				}
	}
	uint32_t* pb = part != nullptr ? part + 6 * (int64_t)blockIdx.x : nullptr;
	block_minmax(mnx, mxx, any, &counters[CT_UX_MIN], &counters[CT_UX_MAX], pb);
	block_minmax(mny, mxy, any, &counters[CT_UY_MIN], &counters[CT_UY_MAX], pb ? pb + 2 : nullptr);
	block_minmax(mnz, mxz, any, &counters[CT_UZ_MIN], &counters[CT_UZ_MAX], pb ? pb + 4 : nullptr);
}

// Sort keys: the occupied box of cell fields (and the sub-world, most
// significant) mapped to a dense integer range, so the radix sort only
// passes over the bits that vary; order is that of the stored keys.
struct KeyCompress
{
	uint32_t ux0, uy0, uz0;
	uint64_t nx, ny, nz;
	int batched;
};

__global__ void k_key_compress(uint64_t* keys, int64_t n, KeyCompress kc)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	KeyFields f = key_fields(keys[i], kc.batched);
	keys[i] = (((uint64_t)f.s * kc.nx + (f.x - kc.ux0)) * kc.ny + (f.y - kc.uy0)) * kc.nz + (f.z - kc.uz0);
}

__global__ void k_key_expand(uint64_t* keys, int64_t n, KeyCompress kc)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	uint64_t c = keys[i];
	uint64_t cz = c % kc.nz;
	uint64_t r = c / kc.nz;
	uint64_t cy = r % kc.ny;
	r /= kc.ny;
	uint64_t cx = r % kc.nx;
	uint64_t cs = r / kc.nx;
	keys[i] = kc.batched ? ((cs << 45) | ((cx + kc.ux0) << 30) | ((cy + kc.uy0) << 15) | (cz + kc.uz0))
						 : (((cx + kc.ux0) << 42) | ((cy + kc.uy0) << 21) | (cz + kc.uz0));
}

// cell field bounds of existing stored keys (kept stable entries)
__global__ void k_key_bounds(const uint64_t* keys, int64_t n, int batched, uint32_t* counters)
{
	uint32_t mnx = 0xFFFFFFFFu, mxx = 0, mny = 0xFFFFFFFFu, mxy = 0, mnz = 0xFFFFFFFFu, mxz = 0;
	bool any = false;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
	{
		KeyFields f = key_fields(keys[i], batched);
		mnx = min(mnx, f.x);
		mxx = max(mxx, f.x);
		mny = min(mny, f.y);
		mxy = max(mxy, f.y);
		mnz = min(mnz, f.z);
		mxz = max(mxz, f.z);
		any = true;
	}
	block_minmax(mnx, mxx, any, &counters[CT_UX_MIN], &counters[CT_UX_MAX]);
	block_minmax(mny, mxy, any, &counters[CT_UY_MIN], &counters[CT_UY_MAX]);
	block_minmax(mnz, mxz, any, &counters[CT_UZ_MIN], &counters[CT_UZ_MAX]);
}

__global__ void k_mark_cells(DBodies B, int32_t n, uint8_t includeBit, int gated, uint8_t cellBit, const uint32_t* cnt)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	if (gated && !(B.subStable[B.sub[i]] & SUB_REBUILD))
		return; // this sub-world keeps its stable tier
	uint8_t bp = B.bp[i];
	bool in = (bp & includeBit) && cnt[i] > 0;
	B.bp[i] = in ? (uint8_t)(bp | cellBit) : (uint8_t)(bp & ~cellBit);
}

// empty cell bounds (min words all ones, max words zero) before an emit
__global__ void k_bounds_init(uint32_t* counters)
{
	const uint32_t i = threadIdx.x;
	if (i < 6u)
		counters[CT_UX_MIN + i] = (i & 1u) ? 0u : 0xFFFFFFFFu;
}

// compression of keys whose field bounds sit in the counters block
static KeyCompress key_compress_from(const uint32_t* hc, int32_t subCount, int* bitsOut)
{
	KeyCompress kc;
	kc.batched = subCount > 0 ? 1 : 0;
	kc.ux0 = hc[CT_UX_MIN];
	kc.uy0 = hc[CT_UY_MIN];
	kc.uz0 = hc[CT_UZ_MIN];
	kc.nx = (uint64_t)(hc[CT_UX_MAX] - kc.ux0) + 1;
	kc.ny = (uint64_t)(hc[CT_UY_MAX] - kc.uy0) + 1;
	kc.nz = (uint64_t)(hc[CT_UZ_MAX] - kc.uz0) + 1;
	uint64_t ns = subCount > 0 ? (uint64_t)subCount : 1;
	uint64_t range = ns * kc.nx * kc.ny * kc.nz; // <= 2^63 (plain) or 2^18 * 2^45 (batched)
	int bits = 0;
	while (bits < 64 && (range - 1) >> bits)
		++bits;
	*bitsOut = bits;
	return kc;
}

bool sync_counters(World& w)
{
	if (!hip_ok(hipMemcpyAsync(w.hostCounters, w.counters.p, kCountersSize * sizeof(uint32_t), hipMemcpyDeviceToHost,
							   w.stream),
				"counter readback"))
		return false;
	return hip_ok(hipStreamSynchronize(w.stream), "stream sync");
}

bool ensure_prim_scratch(World& w, int64_t n)
{
	return w.primScratch.ensure(sort_scratch_bytes(n) + 1024) &&
		   w.scanStatus.ensure((size_t)std::max(scan_status_words(n), sort_status_words(n)));
}

bool build_tier(World& w, int tier)
{
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	uint8_t includeBit = tier == 0 ? BP_ACTIVE : BP_STABLE;
	uint8_t cellBit = tier == 0 ? BP_IN_ACTIVE_CELLS : BP_IN_STABLE_CELLS;
	DBuf<uint64_t>& keys = tier == 0 ? w.activeKeys : w.stableKeys;
	DBuf<uint32_t>& vals = tier == 0 ? w.activeBodies : w.stableBodies;
	std::vector<int32_t>& over = tier == 0 ? w.activeOversized : w.stableOversized;
	DBuf<int32_t>& dOver = tier == 0 ? w.dActiveOversized : w.dStableOversized;
	int64_t& count = tier == 0 ? w.activeCount : w.stableCount;
	DBodies B = body_view(w);
	if (n == 0)
	{
		count = 0;
		over.clear();
		return true;
	}
	// Entries go straight into the tier's buffers (room for at least four
	// per body, or what an earlier build needed), so one counter readback
	// returns the entry count, the oversized count and the cell bounds
	// together; a tier with more entries than room is emitted once more.
	size_t target = std::max<size_t>(std::min(keys.cap, vals.cap), 4 * (size_t)n + 1024);
	if (!w.u32a.ensure(n) || !w.u32b.ensure(n) || !w.u32c.ensure(n) || !w.u32d.ensure(n) || !w.i32a.ensure(n) ||
		!ensure_prim_scratch(w, n) || !keys.ensure(target) || !vals.ensure(target) ||
		!w.boundParts.ensure(6 * (size_t)reduce_grid(n, kBlock)))
		return false;
	int64_t room = (int64_t)std::min(keys.cap, vals.cap);
	uint32_t* cnt = w.u32a.p;
	uint32_t* offs = w.u32b.p;
	uint32_t* ov = w.u32c.p;
	uint32_t* ovOff = w.u32d.p;
	k_tier_count<<<grid_for(n), kBlock, 0, s>>>(B, n, includeBit, 0, w.def.gridCellSize, cnt, ov);
	// entry offsets, and the oversized list in body order
	scan_exclusive2(cnt, offs, w.counters.p + CT_ENTRIES, ov, ovOff, w.counters.p + CT_OVERSIZED, n, w.scanStatus.p, s);
	k_compact_indices<<<grid_for(n), kBlock, 0, s>>>(ov, ovOff, n, w.i32a.p);
	k_mark_cells<<<grid_for(n), kBlock, 0, s>>>(B, n, includeBit, 0, cellBit, cnt);
	const unsigned emitGrid = reduce_grid(n, kBlock);
	k_tier_emit<<<emitGrid, kBlock, 0, s>>>(B, n, includeBit, 0, w.def.gridCellSize, cnt, offs, keys.p, vals.p, room,
											w.counters.p, w.boundParts.p);
	k_bounds_reduce<<<1, 256, 0, s>>>(w.boundParts.p, (int32_t)emitGrid, w.counters.p);
	int32_t* head = (int32_t*)w.hostCounters + kCountersSize + tier * kOverHead;
	const int32_t nHead = std::min(n, kOverHead);
	if (!hip_ok(hipMemcpyAsync(head, w.i32a.p, (size_t)nHead * sizeof(int32_t), hipMemcpyDeviceToHost, s),
				"oversized ids") ||
		!sync_counters(w))
		return false;
	int64_t E = w.hostCounters[CT_ENTRIES];
	int32_t nov = (int32_t)w.hostCounters[CT_OVERSIZED];
	over.assign(head, head + std::min(nov, nHead));
	if (E > room || nov > nHead)
	{
		if (E > room)
		{
			// the same entries again (their cell bounds are min / max over
			// the same keys, so the counters keep their values)
			if (!keys.ensure(E) || !vals.ensure(E))
				return false;
			k_tier_emit<<<emitGrid, kBlock, 0, s>>>(B, n, includeBit, 0, w.def.gridCellSize, cnt, offs, keys.p, vals.p,
													E, w.counters.p, w.boundParts.p);
			k_bounds_reduce<<<1, 256, 0, s>>>(w.boundParts.p, (int32_t)emitGrid, w.counters.p);
		}
		if (nov > nHead)
		{
			over.resize(nov);
			if (!hip_ok(hipMemcpyAsync(over.data() + nHead, w.i32a.p + nHead, (size_t)(nov - nHead) * sizeof(int32_t),
									   hipMemcpyDeviceToHost, s),
						"oversized ids"))
				return false;
		}
		if (!sync_counters(w))
			return false;
	}
	if (nov > 0)
	{
		if (!dOver.ensure(nov))
			return false;
		hipMemcpyAsync(dOver.p, w.i32a.p, nov * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
	}
	count = E;
	if (E == 0)
		return true;
	if (!w.sortKeysAlt.ensure(E) || !w.sortValsAlt.ensure(E) || !ensure_prim_scratch(w, E))
		return false;
	int bits = 0;
	KeyCompress kc = key_compress_from(w.hostCounters, w.subCount, &bits);
	k_key_compress<<<grid_for(E), kBlock, 0, s>>>(keys.p, E, kc);
	// an odd pass count leaves the tier in the ping-pong buffers, which then
	// become the tier's (no copy back)
	if (radix_sort_pairs_in_place_or_alt(keys.p, vals.p, w.sortKeysAlt.p, w.sortValsAlt.p, E, bits, w.primScratch.p,
										 w.scanStatus.p, s))
	{
		std::swap(keys, w.sortKeysAlt);
		std::swap(vals, w.sortValsAlt);
	}
	k_key_expand<<<grid_for(E), kBlock, 0, s>>>(keys.p, E, kc);
	return true;
}

// ------------------------------------------------------------ hashing

constexpr uint64_t kEmptyKey = ~0ull;

MX_HD uint32_t hash_key(uint64_t k)
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return (uint32_t)k;
}

__device__ __forceinline__ uint32_t hash_insert(uint64_t* table, uint32_t mask, uint64_t key)
{
	uint32_t h = hash_key(key) & mask;
	for (;;)
	{
		uint64_t prev = atomicCAS((unsigned long long*)&table[h], (unsigned long long)kEmptyKey,
								  (unsigned long long)key);
		if (prev == kEmptyKey || prev == key)
			return h;
		h = (h + 1) & mask;
	}
}

__device__ __forceinline__ int64_t hash_find(const uint64_t* table, uint32_t mask, uint64_t key)
{
	uint32_t h = hash_key(key) & mask;
	for (;;)
	{
		uint64_t k = table[h];
		if (k == key)
			return h;
		if (k == kEmptyKey)
			return -1;
		h = (h + 1) & mask;
	}
}

__global__ void k_hash_contacts(const int2* bodies, int32_t C, uint64_t* table, uint32_t mask)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	int2 b = bodies[c];
	hash_insert(table, mask, contact_key(b.x, b.y));
}

static uint32_t table_size_for(int64_t n)
{
	uint32_t s = 1024;
	while ((int64_t)s < 2 * n + 64)
		s <<= 1;
	return s;
}

// ------------------------------------------------------------ pairs

struct PairCtx
{
	DBodies B;
	int32_t n;
	float cellSize;
	const uint64_t* act;
	const uint32_t* actBody;
	int64_t nAct;
	const uint64_t* stab;
	const uint32_t* stabBody;
	int64_t nStab;
	const uint64_t* contactTable;
	uint32_t contactMask;
	const uint64_t* jointed;
	int32_t nJointed;
};

// Read-only admission filter; src/world.cpp:1486-1538.
__device__ bool pair_admissible(const PairCtx& P, int32_t a, int32_t b)
{
	if (a == b)
		return false;
	const DBodies& B = P.B;
	if (!(B.bp[a] & BP_MOVED) && !(B.bp[b] & BP_MOVED))
		return false;
	int32_t lo = a < b ? a : b;
	int32_t hi = a < b ? b : a;
	uint32_t fl = B.flags[lo], fh = B.flags[hi];
	if (!(fl & BF_IN_USE) || !(fh & BF_IN_USE))
		return false;
	if (bf_type(fl) != M3D_BODY_DYNAMIC && bf_type(fh) != M3D_BODY_DYNAMIC)
		return false;
	if (!(fl & BF_AWAKE) && !(fh & BF_AWAKE))
		return false;
	float4 lL = B.fatLo[lo], lH = B.fatHi[lo], hL = B.fatLo[hi], hH = B.fatHi[hi];
	if (!(lL.x <= hH.x && lL.y <= hH.y && lL.z <= hH.z && hL.x <= lH.x && hL.y <= lH.y && hL.z <= lH.z))
		return false;
	uint4 fa = B.filter[lo], fb = B.filter[hi];
	int32_t ga = (int32_t)fa.z, gb = (int32_t)fb.z;
	if (ga != 0 && ga == gb)
	{
		if (ga < 0)
			return false;
	}
	else if ((fa.x & fb.y) == 0 || (fb.x & fa.y) == 0)
	{
		return false;
	}
	uint64_t key = contact_key(lo, hi);
	if (hash_find(P.contactTable, P.contactMask, key) >= 0)
		return false;
	if (P.nJointed > 0)
	{
		int32_t l = 0, r = P.nJointed;
		while (l < r)
		{
			int32_t m = (l + r) >> 1;
			if (P.jointed[m] < key)
				l = m + 1;
			else
				r = m;
		}
		if (l < P.nJointed && P.jointed[l] == key)
			return false;
	}
	return true;
}

__device__ __forceinline__ bool dedupe_ok(const PairCtx& P, int32_t a, int32_t b, uint64_t runKey)
{
	float4 la = P.B.fatLo[a], lb = P.B.fatLo[b];
	m3d_vec3 lo2 = m3d_max(f4_v3(la), f4_v3(lb));
	float inv = 1.0f / P.cellSize;
	return body_cell_key(P.B, a, cell_coord(lo2.x, inv), cell_coord(lo2.y, inv), cell_coord(lo2.z, inv)) == runKey;
}

// Candidate pairs in the reference's order (src/world.cpp:1600-1637): cell
// runs of the sorted active tier in key order; per run, active x active
// pairs (i, j > i), then for each stable entry of the same cell (in stable
// order) the pairs with every active entry of the run. That order is cut
// into items, one lane each:
//   A item  an active entry i: its pairs (i, j) for j after i in the run;
//   B item  a stable entry s:  its pairs (i, s) for every i of the run.
// Item t of run r sits at runBase[r] + [A items of r, then B items of r], so
// an exclusive scan of per-item counts gives every item its output range.
struct RunInfo
{
	const uint32_t* flag;   // active entry starts a run
	const uint32_t* runEx;  // exclusive scan of flag: run index at a start
	uint32_t* runStart;
	uint32_t* runEnd;
	uint32_t* runStabLo;    // first stable entry of the run's cell
	uint32_t* runItems;     // A + B items per run (zero beyond the last run)
	const uint32_t* runBase;
};

__global__ void k_run_flags(const uint64_t* act, int64_t nAct, uint32_t* flag)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < nAct)
		flag[i] = (i == 0 || act[i] != act[i - 1]) ? 1u : 0u;
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* a, int64_t n, uint64_t key)
{
	int64_t l = 0, r = n;
	while (l < r)
	{
		int64_t m = (l + r) >> 1;
		if (a[m] < key)
			l = m + 1;
		else
			r = m;
	}
	return l;
}

__global__ void k_run_info(PairCtx P, RunInfo R)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P.nAct || !R.flag[i])
		return;
	uint64_t key = P.act[i];
	int64_t e = i + 1;
	while (e < P.nAct && P.act[e] == key)
		++e;
	int64_t l = lower_bound_u64(P.stab, P.nStab, key);
	int64_t u = l;
	while (u < P.nStab && P.stab[u] == key)
		++u;
	uint32_t r = R.runEx[i];
	R.runStart[r] = (uint32_t)i;
	R.runEnd[r] = (uint32_t)e;
	R.runStabLo[r] = (uint32_t)l;
	R.runItems[r] = (uint32_t)((e - i) + (u - l));
}

// A items: one lane per active entry. emit=0 writes the count at cnt[t],
// emit=1 the keys from out[offs[t]].
__global__ void k_pairs_a(PairCtx P, RunInfo R, int emit, uint32_t* cnt, const uint32_t* offs, uint64_t* out)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P.nAct)
		return;
	uint32_t r = R.runEx[i] + R.flag[i] - 1u;
	int64_t st = R.runStart[r], en = R.runEnd[r];
	int64_t t = (int64_t)R.runBase[r] + (i - st);
	uint64_t key = P.act[i];
	int32_t a = (int32_t)P.actBody[i];
	uint32_t k = 0;
	uint64_t o = emit ? offs[t] : 0;
	for (int64_t y = i + 1; y < en; ++y)
	{
		int32_t b = (int32_t)P.actBody[y];
		if (dedupe_ok(P, a, b, key) && pair_admissible(P, a, b))
		{
			if (emit)
				out[o + k] = contact_key(a < b ? a : b, a < b ? b : a);
			++k;
		}
	}
	if (!emit)
		cnt[t] = k;
}

// B items: one lane per stable entry whose cell has an active run.
__global__ void k_pairs_b(PairCtx P, RunInfo R, int emit, uint32_t* cnt, const uint32_t* offs, uint64_t* out)
{
	int64_t sidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (sidx >= P.nStab)
		return;
	uint64_t key = P.stab[sidx];
	int64_t h = lower_bound_u64(P.act, P.nAct, key);
	if (h >= P.nAct || P.act[h] != key)
		return;
	uint32_t r = R.runEx[h];
	int64_t st = R.runStart[r], en = R.runEnd[r];
	int64_t t = (int64_t)R.runBase[r] + (en - st) + (sidx - (int64_t)R.runStabLo[r]);
	int32_t b = (int32_t)P.stabBody[sidx];
	uint32_t k = 0;
	uint64_t o = emit ? offs[t] : 0;
	for (int64_t x = st; x < en; ++x)
	{
		int32_t a = (int32_t)P.actBody[x];
		if (a == b)
			continue; // stale stable entry of a body that woke
		if (dedupe_ok(P, a, b, key) && pair_admissible(P, a, b))
		{
			if (emit)
				out[o + k] = contact_key(a < b ? a : b, a < b ? b : a);
			++k;
		}
	}
	if (!emit)
		cnt[t] = k;
}

// Coarse-list candidate segments, laid out back to back in one index space.
struct OverSeg
{
	int32_t a;
	int32_t kind;  // 0: all bodies in either tier's cells, 1: list, 2: bodies in active cells
	int32_t list;  // 0 active oversized list, 1 stable oversized list
	int32_t first; // first list index (kind 1) or first body (kinds 0, 2)
	int64_t start; // first global candidate index
	int64_t len;
};

struct OverCtx
{
	const OverSeg* seg; // ascending start
	int32_t nseg;
	const int32_t* actOver;
	const int32_t* stabOver;
};

__device__ bool over_candidate(const PairCtx& P, const OverCtx& O, int64_t g, uint64_t* keyOut)
{
	int32_t lo = 0, hi = O.nseg - 1; // last segment starting at or before g
	while (lo < hi)
	{
		int32_t m = (lo + hi + 1) >> 1;
		if (O.seg[m].start <= g)
			lo = m;
		else
			hi = m - 1;
	}
	const OverSeg S = O.seg[lo];
	int64_t k = g - S.start;
	int32_t a = S.a, b;
	if (S.kind == 1)
	{
		b = (S.list == 0 ? O.actOver : O.stabOver)[S.first + k];
		if (S.list == 1 && b == a)
			return false;
	}
	else
	{
		b = S.first + (int32_t)k;
		uint8_t bp = P.B.bp[b];
		if (S.kind == 0)
		{
			if (b == a || !(bp & (BP_IN_ACTIVE_CELLS | BP_IN_STABLE_CELLS)))
				return false;
		}
		else if (!(bp & BP_IN_ACTIVE_CELLS))
		{
			return false;
		}
	}
	if (!pair_admissible(P, a, b))
		return false;
	*keyOut = contact_key(a < b ? a : b, a < b ? b : a);
	return true;
}

__global__ void k_over_flags(PairCtx P, OverCtx O, int64_t total, uint32_t* flags)
{
	int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= total)
		return;
	uint64_t key;
	flags[g] = over_candidate(P, O, g, &key) ? 1u : 0u;
}

__global__ void k_over_emit(PairCtx P, OverCtx O, int64_t total, const uint32_t* flags, const uint32_t* offs,
							uint64_t* out)
{
	int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= total || !flags[g])
		return;
	uint64_t key;
	over_candidate(P, O, g, &key);
	out[offs[g]] = key;
}

__global__ void k_first_insert(const uint64_t* cand, int64_t n, uint64_t* table, uint32_t* pos, uint32_t mask)
{
	int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	uint32_t h = hash_insert(table, mask, cand[k]);
	atomicMin(&pos[h], (uint32_t)k);
}

__global__ void k_first_flag(const uint64_t* cand, int64_t n, const uint64_t* table, const uint32_t* pos,
							 uint32_t mask, uint32_t* flags)
{
	int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	int64_t h = hash_find(table, mask, cand[k]);
	flags[k] = (h >= 0 && pos[h] == (uint32_t)k) ? 1u : 0u;
}

__global__ void k_compact_u64(const uint64_t* in, const uint32_t* flags, const uint32_t* offs, int64_t n,
							  uint64_t* out)
{
	int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n && flags[k])
		out[offs[k]] = in[k];
}

// Contact initialization; src/world.cpp:1540-1568 (value-initialized
// manifold, combined materials, curved flag).
__global__ void k_create(DBodies B, DContacts C, const uint64_t* keys, int32_t count, int32_t first,
						 const int32_t* dst)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= count)
		return;
	uint64_t key = keys[k];
	int32_t lo = (int32_t)(key >> 32), hi = (int32_t)(uint32_t)key;
	int32_t c = dst != nullptr ? dst[k] : first + k;
	float4 ml = B.mat[lo], mh = B.mat[hi];
	uint32_t fl = B.flags[lo], fh = B.flags[hi];
	float friction = sqrtf(ml.y * mh.y);
	float restitution = ml.z > mh.z ? ml.z : mh.z;
	float rolling = sqrtf(ml.w * mh.w);
	uint32_t cf = 0;
	if ((fl & BF_SENSOR) || (fh & BF_SENSOR))
		cf |= CF_SENSOR;
	if (bf_shape(fl) != M3D_SHAPE_BOX || bf_shape(fh) != M3D_SHAPE_BOX)
		cf |= CF_CURVED;
	C.bodies[c] = make_int2(lo, hi);
	C.hint[c] = make_int2(0, 0);
	C.mat[c] = make_float4(friction, restitution, rolling, u2f(cf));
	C.normal[c] = make_float4(0.0f, 1.0f, 0.0f, i2f(0));
	float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	float4* kc = cache_of(C, c);
	kc[0] = z;
	kc[1] = z;
	kc[2] = z;
	kc[3] = z;
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = pidx(C, c, j);
		C.pA[p] = z;
		C.pB[p] = z;
		C.pJA[p] = z;
		C.pJB[p] = z;
		C.pRA[p] = z;
		C.pRB[p] = z;
		C.pFA[p] = z;
		C.pFB[p] = z;
		C.pSlip[p] = z;
	}
}

static bool ensure_pinned(char*& p, size_t& cap, size_t bytes, const char* what)
{
	if (p != nullptr && cap >= bytes)
		return true;
	if (p != nullptr)
		(void)hipHostFree(p);
	p = nullptr;
	cap = bytes * 3 / 2 + 4096;
	if (!hip_ok(hipHostMalloc((void**)&p, cap, hipHostMallocDefault), what))
	{
		cap = 0;
		return false;
	}
	return true;
}

bool enumerate_and_create(World& w)
{
	w.newCount = 0;
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	int32_t C0 = w.contactCount;
	// existing-contact map
	uint32_t tsz = table_size_for(C0);
	if (!w.hashKeys.ensure(tsz) || !w.hashVals.ensure(tsz))
		return false;
	launch_fill_u64(w.hashKeys.p, kEmptyKey, tsz, s);
	if (C0 > 0)
		k_hash_contacts<<<grid_for(C0), kBlock, 0, s>>>(w.cc.bodies.p, C0, w.hashKeys.p, tsz - 1);

	PairCtx P;
	P.B = body_view(w);
	P.n = n;
	P.cellSize = w.def.gridCellSize;
	P.act = w.activeKeys.p;
	P.actBody = w.activeBodies.p;
	P.nAct = w.activeCount;
	P.stab = w.stableKeys.p;
	P.stabBody = w.stableBodies.p;
	P.nStab = w.stableCount;
	P.contactTable = w.hashKeys.p;
	P.contactMask = tsz - 1;
	P.jointed = w.jointedKeys.p;
	P.nJointed = w.jointedCount;

	int64_t E = w.activeCount;
	int64_t SB = w.stableCount;
	int64_t T = E + SB; // grid items: at most one per active entry plus one per stable entry
	// Coarse-list segments in the reference's serial order
	// (src/world.cpp:1643-1702). Batched: each oversized body meets only its
	// own sub-world's bodies and oversized lists, which keeps every
	// sub-world's candidate order that of a world of its own.
	OverCtx O;
	O.actOver = w.dActiveOversized.p;
	O.stabOver = w.dStableOversized.p;
	int64_t ovTotal = 0;
	std::vector<OverSeg> segs;
	auto addSeg = [&](int32_t a, int kind, int list, int first, int64_t len) {
		if (len <= 0)
			return;
		OverSeg S;
		S.a = a;
		S.kind = kind;
		S.list = list;
		S.first = first;
		S.start = ovTotal;
		S.len = len;
		segs.push_back(S);
		ovTotal += len;
	};
	const std::vector<int32_t>& ao = w.activeOversized;
	const std::vector<int32_t>& so = w.stableOversized;
	if (w.subCount == 0)
	{
		for (size_t oi = 0; oi < ao.size(); ++oi)
		{
			addSeg(ao[oi], 0, 0, 0, n);
			addSeg(ao[oi], 1, 0, (int)oi + 1, (int64_t)ao.size() - (int64_t)oi - 1);
			addSeg(ao[oi], 1, 1, 0, (int64_t)so.size());
		}
		for (size_t si = 0; si < so.size(); ++si)
			addSeg(so[si], 2, 0, 0, n);
	}
	else
	{
		// both lists ascend by body, so a sub-world's entries are one range
		auto subRange = [&](const std::vector<int32_t>& l, int32_t sw, size_t* b, size_t* e) {
			int32_t lo = w.subBodyStart[sw], hi = w.subBodyStart[sw + 1];
			*b = (size_t)(std::lower_bound(l.begin(), l.end(), lo) - l.begin());
			*e = (size_t)(std::lower_bound(l.begin(), l.end(), hi) - l.begin());
		};
		for (size_t oi = 0; oi < ao.size(); ++oi)
		{
			int32_t sw = w.bodySubHost[ao[oi]];
			int32_t b0 = w.subBodyStart[sw], nb = w.subBodyStart[sw + 1] - b0;
			size_t ab, ae, sb, se;
			subRange(ao, sw, &ab, &ae);
			subRange(so, sw, &sb, &se);
			addSeg(ao[oi], 0, 0, b0, nb);
			addSeg(ao[oi], 1, 0, (int)oi + 1, (int64_t)ae - (int64_t)oi - 1);
			addSeg(ao[oi], 1, 1, (int)sb, (int64_t)se - (int64_t)sb);
		}
		for (size_t si = 0; si < so.size(); ++si)
		{
			int32_t sw = w.bodySubHost[so[si]];
			int32_t b0 = w.subBodyStart[sw], nb = w.subBodyStart[sw + 1] - b0;
			addSeg(so[si], 2, 0, b0, nb);
		}
	}
	O.nseg = (int32_t)segs.size();
	if (!segs.empty())
	{
		size_t words = (segs.size() * sizeof(OverSeg) + 7) / 8;
		if (!w.overSegs.ensure(words))
			return false;
		// pinned, so the upload does not wait for the stream (the earlier
		// step's upload from the same memory has ended at its readbacks)
		if (!ensure_pinned(w.pinnedSegs, w.pinnedSegsCap, words * 8, "coarse-list segments"))
			return false;
		memcpy(w.pinnedSegs, segs.data(), segs.size() * sizeof(OverSeg));
		if (!hip_ok(hipMemcpyAsync(w.overSegs.p, w.pinnedSegs, words * 8, hipMemcpyHostToDevice, s),
					"coarse-list segments"))
			return false;
		O.seg = (const OverSeg*)w.overSegs.p;
	}
	else
	{
		O.seg = nullptr;
	}
	// Both candidate counts (grid items and admitted coarse-list pairs) come
	// back with one counter readback.
	if (E > 0 && (!w.pairRuns.ensure(7 * (size_t)E + 2 * (size_t)T)))
		return false;
	if (ovTotal > 0 && (!w.u32c.ensure(ovTotal) || !w.u32d.ensure(ovTotal)))
		return false;
	if ((E > 0 || ovTotal > 0) && !ensure_prim_scratch(w, std::max(E > 0 ? T : 0, ovTotal)))
		return false;
	RunInfo R;
	uint32_t* itemCnt = nullptr;
	uint32_t* itemOffs = nullptr;
	if (E > 0)
	{
		uint32_t* base = w.pairRuns.p;
		uint32_t* flag = base;
		uint32_t* runEx = base + E;
		R.flag = flag;
		R.runEx = runEx;
		R.runStart = base + 2 * E;
		R.runEnd = base + 3 * E;
		R.runStabLo = base + 4 * E;
		R.runItems = base + 5 * E;
		uint32_t* runBase = base + 6 * E;
		R.runBase = runBase;
		itemCnt = base + 7 * E;
		itemOffs = itemCnt + T;
		FillSegs F;
		F.add(R.runItems, 0u, E);
		F.add(itemCnt, 0u, T);
		launch_fill_multi(F, s);
		k_run_flags<<<grid_for(E), kBlock, 0, s>>>(w.activeKeys.p, E, flag);
		scan_exclusive(flag, runEx, E, nullptr, w.scanStatus.p, s);
		k_run_info<<<grid_for(E), kBlock, 0, s>>>(P, R);
		scan_exclusive(R.runItems, runBase, E, nullptr, w.scanStatus.p, s);
		w.mark_begin(M3DX_K_PAIRS);
		k_pairs_a<<<grid_for(E), kBlock, 0, s>>>(P, R, 0, itemCnt, nullptr, nullptr);
		if (SB > 0)
			k_pairs_b<<<grid_for(SB), kBlock, 0, s>>>(P, R, 0, itemCnt, nullptr, nullptr);
		w.mark_end();
		scan_exclusive(itemCnt, itemOffs, T, w.counters.p + CT_CANDIDATES, w.scanStatus.p, s);
	}
	uint32_t* ovFlags = nullptr;
	uint32_t* ovOffs = nullptr;
	if (ovTotal > 0)
	{
		ovFlags = w.u32c.p;
		ovOffs = w.u32d.p;
		k_over_flags<<<grid_for(ovTotal), kBlock, 0, s>>>(P, O, ovTotal, ovFlags);
		scan_exclusive(ovFlags, ovOffs, ovTotal, w.counters.p + CT_SEG_COUNT, w.scanStatus.p, s);
	}
	if ((E > 0 || ovTotal > 0) && !sync_counters(w))
		return false;
	int64_t gridTotal = E > 0 ? (int64_t)w.hostCounters[CT_CANDIDATES] : 0;
	int64_t ovAdmitted = ovTotal > 0 ? (int64_t)w.hostCounters[CT_SEG_COUNT] : 0;
	int64_t total = gridTotal + ovAdmitted;
	if (total == 0)
	{
		return true;
	}
	if (!w.u64a.ensure(total) || !w.u64b.ensure(total))
		return false;
	if (gridTotal > 0)
	{
		w.mark_begin(M3DX_K_PAIRS);
		k_pairs_a<<<grid_for(E), kBlock, 0, s>>>(P, R, 1, nullptr, itemOffs, w.u64a.p);
		if (SB > 0)
			k_pairs_b<<<grid_for(SB), kBlock, 0, s>>>(P, R, 1, nullptr, itemOffs, w.u64a.p);
		w.mark_end();
	}
	if (ovAdmitted > 0)
		k_over_emit<<<grid_for(ovTotal), kBlock, 0, s>>>(P, O, ovTotal, ovFlags, ovOffs, w.u64a.p + gridTotal);

	// first admitted occurrence of each key wins
	uint32_t csz = table_size_for(total);
	if (!w.u64c.ensure(csz) || !w.u32a.ensure(std::max<int64_t>(csz, total)) || !w.u32b.ensure(total) ||
		!w.u32c.ensure(total) || !ensure_prim_scratch(w, total))
		return false;
	{
		// the key table (all ones: kEmptyKey) and its first-occurrence slots
		FillSegs F;
		F.add((uint32_t*)w.u64c.p, 0xFFFFFFFFu, 2 * (int64_t)csz);
		F.add(w.u32a.p, 0xFFFFFFFFu, csz);
		launch_fill_multi(F, s);
	}
	k_first_insert<<<grid_for(total), kBlock, 0, s>>>(w.u64a.p, total, w.u64c.p, w.u32a.p, csz - 1);
	k_first_flag<<<grid_for(total), kBlock, 0, s>>>(w.u64a.p, total, w.u64c.p, w.u32a.p, csz - 1, w.u32b.p);
	scan_exclusive(w.u32b.p, w.u32c.p, total, w.counters.p + CT_NEW_CONTACTS, w.scanStatus.p, s);
	k_compact_u64<<<grid_for(total), kBlock, 0, s>>>(w.u64a.p, w.u32b.p, w.u32c.p, total, w.u64b.p);
	if (!sync_counters(w))
		return false;
	int32_t K = (int32_t)w.hostCounters[CT_NEW_CONTACTS];
	if (w.subCount > 0)
		return stage_new_contacts(w, w.u64b.p, K); // placed with the removals
	if (K == 0)
		return true;
	if (!ensure_device_contacts(w, C0 + K))
		return false;
	k_create<<<grid_for(K), kBlock, 0, s>>>(body_view(w), contact_view(w), w.u64b.p, K, C0, nullptr);
	w.contactCount = C0 + K;
	return true;
}

// ------------------------------------------------------------ removal

__global__ void k_remove_flags(DBodies B, const int2* bodies, int32_t C, uint32_t* flags)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	int2 ab = bodies[c];
	uint32_t f = 0;
	if ((B.sb[ab.x] & SB_AWAKE) || (B.sb[ab.y] & SB_AWAKE))
	{
		float4 aL = B.fatLo[ab.x], aH = B.fatHi[ab.x], bL = B.fatLo[ab.y], bH = B.fatHi[ab.y];
		bool ov = aL.x <= bH.x && aL.y <= bH.y && aL.z <= bH.z && bL.x <= aH.x && bL.y <= aH.y && bL.z <= aH.z;
		f = ov ? 0u : 1u;
	}
	flags[c] = f;
}

// Holes (removed below newC = C - *removed), tails (kept at or above it),
// end-event flags, and the removal flags of the tail region [newC, C)
// packed from 0 (tailRm).
__global__ void k_hole_tail_flags(const uint32_t* rm, const float4* mat, int32_t C, const uint32_t* removed,
								  uint32_t* hole, uint32_t* tail, uint32_t* ev, uint32_t* tailRm)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	const int32_t newC = C - (int32_t)*removed;
	uint32_t r = rm[c];
	hole[c] = (r && c < newC) ? 1u : 0u;
	tail[c] = (!r && c >= newC) ? 1u : 0u;
	ev[c] = (r && (f2u(mat[c].w) & CF_TOUCHING)) ? 1u : 0u;
	if (c >= newC)
		tailRm[c - newC] = r;
}

__global__ void k_make_moves(const uint32_t* hole, const uint32_t* holeOff, const uint32_t* tail,
							 const uint32_t* tailOff, int32_t C, int2* moves)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	if (hole[c])
		moves[holeOff[c]].x = c;
	if (tail[c])
		moves[tailOff[c]].y = c;
}

__global__ void k_set_move_sources(int2* moves, const int32_t* src, int32_t H)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k < H)
		moves[k].y = src[k];
}

// Point j of contact s, in every point column, to point j of contact d, and
// a quarter of the contact's own columns: four consecutive lanes move one
// contact, so a wave moves 16 contacts' points as contiguous runs (a
// contact's four points are adjacent in each column).
// Part j of contact s to slot d: point j (unless point == false) and a
// quarter of the contact's own fields.
__device__ __forceinline__ void copy_contact_part(const DContacts& S, int32_t s, const DContacts& D, int32_t d, int j,
												  bool point = true)
{
	if (point)
	{
		int32_t pd = pidx(D, d, j), ps = pidx(S, s, j);
		D.pA[pd] = S.pA[ps];
		D.pB[pd] = S.pB[ps];
		D.pJA[pd] = S.pJA[ps];
		D.pJB[pd] = S.pJB[ps];
		D.pRA[pd] = S.pRA[ps];
		D.pRB[pd] = S.pRB[ps];
		D.pFA[pd] = S.pFA[ps];
		D.pFB[pd] = S.pFB[ps];
		D.pSlip[pd] = S.pSlip[ps];
	}
	cache_of(D, d)[j] = cache_of(S, s)[j];
	switch (j)
	{
	case 0:
		D.bodies[d] = S.bodies[s];
		D.hint[d] = S.hint[s];
		break;
	case 1:
		D.mat[d] = S.mat[s];
		break;
	case 2:
		D.normal[d] = S.normal[s];
		break;
	default:
		break;
	}
}

// One lane per (move, point).
__global__ void k_apply_moves(DContacts C, const int2* moves, int32_t count)
{
	int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	int32_t k = (int32_t)(t >> 2);
	if (k >= count)
		return;
	copy_contact_part(C, moves[k].y, C, moves[k].x, (int)(t & 3));
}

bool apply_moves(World& w, const int2* dMoves, int32_t count)
{
	if (count > 0)
		k_apply_moves<<<grid_for(4 * (int64_t)count), kBlock, 0, w.stream>>>(contact_view(w), dMoves, count);
	return true;
}

// (body a, body b, sensor, contact) of each flagged contact, at its offset
__global__ void k_gather_flagged_events(const int2* bodies, const float4* mat, const uint32_t* flag,
										const uint32_t* offs, int32_t C, int4* out)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C || !flag[c])
		return;
	int2 b = bodies[c];
	uint32_t cf = f2u(mat[c].w);
	out[offs[c]] = make_int4(b.x, b.y, (cf & CF_SENSOR) ? 1 : 0, c);
}

bool remove_separated(World& w)
{
	hipStream_t s = w.stream;
	int32_t C = w.contactCount;
	if (C == 0)
		return true;
	if (!w.u32a.ensure(C) || !w.u32b.ensure(C) || !w.u32c.ensure(C) || !w.u32d.ensure(C) || !w.i32a.ensure(C) ||
		!w.i32b.ensure(C) || !w.i32c.ensure(C) || !w.i32d.ensure(C) ||
		!w.u64a.ensure((size_t)C * 2) || !ensure_prim_scratch(w, C))
		return false;
	// One readback returns the removed count, the hole / tail / end-event
	// counts, the end events and the removal flags of the tail region when
	// at most `spec` contacts go (spec follows twice the last count); more
	// take a second copy. The pinned block is reused by the next call only
	// after that call's readback, which follows every copy issued here.
	const int32_t spec = std::min(C, std::max<int32_t>(4096, w.removalSpec));
	size_t specBytes = (size_t)spec * (sizeof(int4) + 2 * sizeof(uint32_t));
	if (!ensure_pinned(w.pinnedRemoval, w.pinnedRemovalCap, specBytes, "removal staging"))
		return false;
	uint32_t* rm = w.u32a.p;
	uint32_t* hole = w.u32b.p;
	uint32_t* tail = w.u32c.p;
	uint32_t* ev = w.u32d.p;
	uint32_t* tailRm = (uint32_t*)w.i32d.p;
	uint32_t* holeOff = (uint32_t*)w.i32a.p;
	uint32_t* tailOff = (uint32_t*)w.i32b.p;
	uint32_t* evOff = (uint32_t*)w.i32c.p;
	k_remove_flags<<<grid_for(C), kBlock, 0, s>>>(body_view(w), w.cc.bodies.p, C, rm);
	scan_exclusive(rm, hole, C, w.counters.p + CT_REMOVED, w.scanStatus.p, s); // offsets unused: the count
	k_hole_tail_flags<<<grid_for(C), kBlock, 0, s>>>(rm, w.cc.mat.p, C, w.counters.p + CT_REMOVED, hole, tail, ev,
													 tailRm);
	scan_exclusive3(hole, holeOff, w.counters.p + CT_HOLES, tail, tailOff, w.counters.p + CT_TAILS, ev, evOff,
					w.counters.p + CT_END_EVENTS, C, w.scanStatus.p, s);
	// end events of touching removed pairs, ascending contact order (the
	// host reads them back to front)
	k_gather_flagged_events<<<grid_for(C), kBlock, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, ev, evOff, C, (int4*)w.u64a.p);
	int4* evHost = (int4*)w.pinnedRemoval;
	uint32_t* flags = (uint32_t*)(w.pinnedRemoval + (size_t)spec * sizeof(int4));
	int32_t* order = (int32_t*)(w.pinnedRemoval + (size_t)spec * (sizeof(int4) + sizeof(uint32_t)));
	if (!hip_ok(hipMemcpyAsync(evHost, w.u64a.p, (size_t)spec * sizeof(int4), hipMemcpyDeviceToHost, s),
				"end events") ||
		!hip_ok(hipMemcpyAsync(flags, tailRm, (size_t)spec * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
				"removal flags") ||
		!sync_counters(w))
		return false;
	int32_t R = (int32_t)w.hostCounters[CT_REMOVED];
	w.removalSpec = 2 * R;
	if (R == 0)
		return true;
	int32_t newC = C - R;
	int32_t H = (int32_t)w.hostCounters[CT_HOLES];
	int32_t NE = (int32_t)w.hostCounters[CT_END_EVENTS];
	if (H != (int32_t)w.hostCounters[CT_TAILS] || NE > R)
	{
		set_error("removal bookkeeping mismatch (%d holes, %u tails, %d end events)", H, w.hostCounters[CT_TAILS], NE);
		return false;
	}
	if (R > spec)
	{
		// a larger block (its contents are not kept) and both lists again
		if (!ensure_pinned(w.pinnedRemoval, w.pinnedRemovalCap, (size_t)R * (sizeof(int4) + 2 * sizeof(uint32_t)),
						   "removal staging"))
			return false;
		evHost = (int4*)w.pinnedRemoval;
		flags = (uint32_t*)(w.pinnedRemoval + (size_t)R * sizeof(int4));
		order = (int32_t*)(w.pinnedRemoval + (size_t)R * (sizeof(int4) + sizeof(uint32_t)));
		if (!hip_ok(hipMemcpyAsync(evHost, w.u64a.p, (size_t)NE * sizeof(int4), hipMemcpyDeviceToHost, s),
					"end events") ||
			!hip_ok(hipMemcpyAsync(flags, tailRm, (size_t)R * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
					"removal flags") ||
			!hip_ok(hipStreamSynchronize(s), "removal staging"))
			return false;
	}
	if (H > 0)
	{
		// The reference removes in descending order, moving the current last
		// contact into each removed slot (src/world.cpp:1722-1734, 100-119).
		// The removed slots at or beyond newC come first and may take in a
		// contact from the end too, so the survivors of [newC, C) stand in the
		// order a replay of that region (R slots) leaves them; the holes below
		// newC then take them from the top down: the k-th hole (ascending)
		// receives the k-th of them.
		if (!w.moves.ensure(H) || !w.tailOrder.ensure(R))
			return false;
		for (int32_t i = 0; i < R; ++i)
			order[i] = newC + i;
		int32_t top = R - 1;
		for (int32_t r = R - 1; r >= 0; --r)
			if (flags[r])
			{
				if (r != top)
					order[r] = order[top];
				--top;
			}
		if (top + 1 != H)
		{
			set_error("removal replay mismatch (%d survivors, %d holes)", top + 1, H);
			return false;
		}
		if (!hip_ok(hipMemcpyAsync(w.tailOrder.p, order, (size_t)H * sizeof(int32_t), hipMemcpyHostToDevice, s),
					"removal order"))
			return false;
		k_make_moves<<<grid_for(C), kBlock, 0, s>>>(hole, holeOff, tail, tailOff, C, w.moves.p);
		k_set_move_sources<<<grid_for(H), kBlock, 0, s>>>(w.moves.p, w.tailOrder.p, H);
		k_apply_moves<<<grid_for(4 * (int64_t)H), kBlock, 0, s>>>(contact_view(w), w.moves.p, H);
	}
	for (int32_t k = NE - 1; k >= 0; --k)
	{
		const int4& e = evHost[k];
		m3d_contact_event ce;
		ce.bodyA = { e.x, w.bodies[e.x].generation };
		ce.bodyB = { e.y, w.bodies[e.y].generation };
		ce.isSensor = e.z != 0;
		w.endEvents.push_back(ce);
	}
	w.contactCount = newC;
	return true;
}

// ------------------------------------------------------------ bulk despawn

// rank[slot] = rank for the destroyed bodies of a bulk despawn (INT32_MAX
// elsewhere: the array is restored after each call)
__global__ void k_resp_rank(int32_t* rank, const int2* destroyed, int32_t count, int32_t clear)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k < count)
		rank[destroyed[k].x] = clear ? INT32_MAX : destroyed[k].y;
}

__global__ void k_resp_flags(const int2* bodies, int32_t C, const int32_t* rank, uint32_t* flag)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	int2 b = bodies[c];
	flag[c] = (rank[b.x] != INT32_MAX || rank[b.y] != INT32_MAX) ? 1u : 0u;
}

__global__ void k_resp_compact(const int2* bodies, const float4* mat, int32_t C, const uint32_t* flag,
							   const uint32_t* offs, int4* out)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C || !flag[c])
		return;
	int2 b = bodies[c];
	out[offs[c]] = make_int4(b.x, b.y, (int32_t)f2u(mat[c].w), c);
}

bool gather_despawn_contacts(World& w, const std::vector<int2>& destroyed, std::vector<int4>& out)
{
	out.clear();
	int32_t C = w.contactCount;
	int32_t n = w.deviceBodyCount;
	int32_t nD = (int32_t)destroyed.size();
	if (C == 0 || nD == 0)
		return true;
	hipStream_t s = w.stream;
	// the rank array lives across calls, all INT32_MAX between them
	if ((int64_t)w.respRank.cap < (int64_t)n)
	{
		if (!w.respRank.ensure(n))
			return false;
		launch_fill_i32(w.respRank.p, INT32_MAX, (int64_t)w.respRank.cap, s);
	}
	if (!w.respList.ensure(nD) || !w.u32a.ensure(C) || !w.u32b.ensure(C) || !w.respOut.ensure(C) ||
		!ensure_prim_scratch(w, C))
		return false;
	constexpr int32_t kFirst = 8192; // removed contacts fetched with the count
	if (w.pinnedResp == nullptr || w.pinnedRespCap < (size_t)std::max(C, kFirst))
	{
		if (w.pinnedResp != nullptr)
			(void)hipHostFree(w.pinnedResp);
		w.pinnedResp = nullptr;
		w.pinnedRespCap = (size_t)std::max(C, kFirst) * 3 / 2;
		if (!hip_ok(hipHostMalloc((void**)&w.pinnedResp, w.pinnedRespCap * sizeof(int4), hipHostMallocDefault),
					"pinned despawn list"))
		{
			w.pinnedRespCap = 0;
			return false;
		}
	}
	if (!hip_ok(hipMemcpyAsync(w.respList.p, destroyed.data(), nD * sizeof(int2), hipMemcpyHostToDevice, s),
				"despawn list"))
		return false;
	k_resp_rank<<<grid_for(nD), kBlock, 0, s>>>(w.respRank.p, w.respList.p, nD, 0);
	k_resp_flags<<<grid_for(C), kBlock, 0, s>>>(w.cc.bodies.p, C, w.respRank.p, w.u32a.p);
	scan_exclusive(w.u32a.p, w.u32b.p, C, w.counters.p + CT_RESPAWN_REMOVED, w.scanStatus.p, s);
	k_resp_compact<<<grid_for(C), kBlock, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, C, w.u32a.p, w.u32b.p, w.respOut.p);
	k_resp_rank<<<grid_for(nD), kBlock, 0, s>>>(w.respRank.p, w.respList.p, nD, 1);
	int32_t first = std::min(C, kFirst);
	if (!hip_ok(hipMemcpyAsync(w.pinnedResp, w.respOut.p, first * sizeof(int4), hipMemcpyDeviceToHost, s),
				"despawn contacts") ||
		!sync_counters(w))
		return false;
	int32_t R = (int32_t)w.hostCounters[CT_RESPAWN_REMOVED];
	if (R > first)
	{
		if (!hip_ok(hipMemcpyAsync(w.pinnedResp + first, w.respOut.p + first, (size_t)(R - first) * sizeof(int4),
								   hipMemcpyDeviceToHost, s),
					"despawn contacts") ||
			!hip_ok(hipStreamSynchronize(s), "despawn contacts"))
			return false;
	}
	out.assign(w.pinnedResp, w.pinnedResp + R);
	return true;
}

// ------------------------------------------------------------ host sync helpers

__global__ void k_material_refresh(DBodies B, DContacts C, int32_t count, const uint8_t* flag)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= count)
		return;
	int2 ab = C.bodies[c];
	if (!flag[ab.x] && !flag[ab.y])
		return;
	float4 ma = B.mat[ab.x], mb = B.mat[ab.y];
	float4 m = C.mat[c];
	m.x = sqrtf(ma.y * mb.y);
	m.y = ma.z > mb.z ? ma.z : mb.z;
	m.z = sqrtf(ma.w * mb.w);
	C.mat[c] = m;
}

// Sidecar record of every contact, and the (body, contact side) pairs the
// per-body contact index is sorted from: entry 2c + side is contact c's
// body A (side 0) or B (side 1). A sort replaces per-body atomic counters,
// which serialize on bodies in many contacts (a ground).
__global__ void k_pack_sidecar(const int2* bodies, const float4* mat, int32_t C, int4* out, uint64_t* keys,
							   uint32_t* vals)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c < C)
	{
		int2 b = bodies[c];
		out[c] = make_int4(b.x, b.y, (int32_t)f2u(mat[c].w), c);
		keys[2 * (int64_t)c] = (uint64_t)(uint32_t)b.x;
		keys[2 * (int64_t)c + 1] = (uint64_t)(uint32_t)b.y;
		vals[2 * (int64_t)c] = (uint32_t)c;
		vals[2 * (int64_t)c + 1] = (uint32_t)c | 0x80000000u;
	}
}

// per-body contact index from the sorted pairs: list[start[b] ..] = body
// b's contacts (ascending), slot[c] = the two entries of contact c
__global__ void k_contact_csr(const uint32_t* vals, int64_t m, int32_t* list, int32_t* slot)
{
	int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= m)
		return;
	uint32_t v = vals[k];
	int32_t c = (int32_t)(v & 0x7FFFFFFFu);
	list[k] = c;
	slot[2 * (int64_t)c + (v >> 31)] = (int32_t)k;
}

// start[b] = first sorted entry of body b (lower bound), b = 0 .. n
__global__ void k_csr_start(const uint64_t* keys, int64_t m, int32_t n, uint32_t* start)
{
	int32_t b = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (b > n)
		return;
	int64_t lo = 0, hi = m;
	while (lo < hi)
	{
		int64_t mid = (lo + hi) >> 1;
		if (keys[mid] < (uint64_t)b)
			lo = mid + 1;
		else
			hi = mid;
	}
	start[b] = (uint32_t)lo;
}

bool prefetch_sidecar(World& w)
{
	int32_t C = w.contactCount;
	w.sidecarPrefetched = false;
	if (C == 0)
		return true;
	if (w.pinnedSidecarCap < (size_t)C)
	{
		if (w.pinnedSidecar != nullptr)
			(void)hipHostFree(w.pinnedSidecar);
		w.pinnedSidecar = nullptr;
		size_t cap = (size_t)C + (size_t)C / 2 + 1024;
		if (hipHostMalloc((void**)&w.pinnedSidecar, cap * sizeof(int4), 0) != hipSuccess)
		{
			w.pinnedSidecarCap = 0;
			return false;
		}
		w.pinnedSidecarCap = cap;
	}
	int32_t n = w.deviceBodyCount;
	size_t csrInts = (size_t)n + 1 + 4 * (size_t)C; // start, list (2C), slot (2C)
	if (w.pinnedCsrCap < csrInts)
	{
		if (w.pinnedCsr != nullptr)
			(void)hipHostFree(w.pinnedCsr);
		w.pinnedCsr = nullptr;
		size_t cap = csrInts + csrInts / 2 + 1024;
		if (hipHostMalloc((void**)&w.pinnedCsr, cap * sizeof(int32_t), 0) != hipSuccess)
		{
			w.pinnedCsrCap = 0;
			return false;
		}
		w.pinnedCsrCap = cap;
	}
	if (w.sidecarReady == nullptr && hipEventCreateWithFlags(&w.sidecarReady, hipEventDisableTiming) != hipSuccess)
		return false;
	// device layout: start (n + 1), list (2C), slot (2C)
	const int64_t m = 2 * (int64_t)C;
	if (!w.dSidecar.ensure(C) || !w.dCsr.ensure(csrInts) || !w.csrKeys.ensure(m) || !w.csrKeysAlt.ensure(m) ||
		!w.csrVals.ensure(m) || !w.csrValsAlt.ensure(m) || !ensure_prim_scratch(w, m))
		return false;
	hipStream_t s = w.stream;
	uint32_t* start = (uint32_t*)w.dCsr.p;
	int32_t* list = w.dCsr.p + n + 1;
	int32_t* slot = list + m;
	int bits = 1;
	while (bits < 31 && ((int64_t)1 << bits) <= (int64_t)n)
		++bits;
	k_pack_sidecar<<<grid_for(C), kBlock, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, C, w.dSidecar.p, w.csrKeys.p,
												  w.csrVals.p);
	radix_sort_pairs(w.csrKeys.p, w.csrVals.p, w.csrKeysAlt.p, w.csrValsAlt.p, m, bits, w.primScratch.p,
					 w.scanStatus.p, s);
	k_csr_start<<<grid_for((int64_t)n + 1), kBlock, 0, s>>>(w.csrKeys.p, m, n, start);
	k_contact_csr<<<grid_for(m), kBlock, 0, s>>>(w.csrVals.p, m, list, slot);
	if (hipMemcpyAsync(w.pinnedSidecar, w.dSidecar.p, (size_t)C * sizeof(int4), hipMemcpyDeviceToHost, s) !=
			hipSuccess ||
		hipMemcpyAsync(w.pinnedCsr, w.dCsr.p, csrInts * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipEventRecord(w.sidecarReady, s) != hipSuccess)
		return false;
	w.sidecarPrefetched = true;
	w.sidecarPrefetchCount = C;
	w.sidecarPrefetchN = n;
	return true;
}

// WakeBodyInternal (src/world.cpp:121-140) for listed slots on the device.
__global__ void k_wake_list(DBodies B, const int32_t* idx, int32_t count, int32_t n)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= count)
		return;
	int32_t i = idx[k];
	if (i < 0 || i >= n || !(B.flags[i] & BF_IN_USE) || bf_type(B.flags[i]) == M3D_BODY_STATIC)
		return;
	B.flags[i] |= BF_AWAKE;
	B.sb[i] |= SB_AWAKE;
	float4 a = B.iw1[i];
	a.w = 0.0f;
	B.iw1[i] = a;
	float4 e = B.iw2[i];
	e.w = 4.0f * kSleepEnergyThresh;
	B.iw2[i] = e;
}

bool launch_wake_list(World& w)
{
	if (w.deviceWake.empty())
		return true;
	w.wakeUpload.swap(w.deviceWake);
	w.deviceWake.clear();
	int32_t count = (int32_t)w.wakeUpload.size();
	if (!w.dWake.ensure(count))
		return false;
	if (!hip_ok(hipMemcpyAsync(w.dWake.p, w.wakeUpload.data(), count * sizeof(int32_t), hipMemcpyHostToDevice,
							   w.stream),
				"wake list"))
		return false;
	k_wake_list<<<grid_for(count), kBlock, 0, w.stream>>>(body_view(w), w.dWake.p, count, w.deviceBodyCount);
	return true;
}

void launch_material_refresh(World& w, const uint8_t* dFlags)
{
	if (w.contactCount > 0)
		k_material_refresh<<<grid_for(w.contactCount), kBlock, 0, w.stream>>>(body_view(w), contact_view(w),
																			   w.contactCount, dFlags);
}

// stage layout: 18 float4 columns back to back, each `count` long, in the
// order pos, rot, vel, angVel, prevPos, prevRot, prevVel, prevAngVel, invIL,
// iw0, iw1, iw2, shape, mat, force, torque, fatLo, fatHi
__global__ void k_upload_scatter(DBodies B, int32_t count, const int32_t* idx, const float4* st, const uint32_t* sf,
								 const uint4* sfil, const uint8_t* sbp)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= count)
		return;
	int32_t i = idx[k];
	float4* cols[18] = { B.pos, B.rot, B.vel, B.angVel, B.prevPos, B.prevRot, B.prevVel, B.prevAngVel, B.invIL,
						 B.iw0, B.iw1, B.iw2, B.shape, B.mat, B.force, B.torque, B.fatLo, B.fatHi };
	for (int c = 0; c < 18; ++c)
		cols[c][i] = st[(int64_t)c * count + k];
	B.flags[i] = sf[k];
	B.filter[i] = sfil[k];
	B.bp[i] = sbp[k];
}

void launch_upload_scatter(World& w, int32_t count, const int32_t* dIdx, const float4* stage, const uint32_t* stageFlags,
						   const uint4* stageFilter, const uint8_t* stageBp)
{
	if (count > 0)
		k_upload_scatter<<<grid_for(count), kBlock, 0, w.stream>>>(body_view(w), count, dIdx, stage, stageFlags,
																	stageFilter, stageBp);
}

// ------------------------------------------------------------ batched sub-worlds
//
// A batch is one device world whose bodies are partitioned into sub-worlds
// (index ranges). Cell keys carry the sub-world, coarse lists stay inside it
// and contacts are kept contiguous per sub-world, each range in the order a
// world of its own would hold; every later stage works per contact, body or
// island, none of which ever spans two sub-worlds.

static int bits_for(uint64_t range)
{
	int b = 0;
	while (b < 64 && (range - 1) >> b)
		++b;
	return b;
}

// flags[t]: SUB_DIRTY and SUB_FORCE bits set by the host since the last step
__global__ void k_sub_or_dirty(uint32_t* subStable, const uint8_t* flags, int32_t S)
{
	int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (t < S && flags[t])
		subStable[t] |= (uint32_t)flags[t] & (SUB_DIRTY | SUB_FORCE);
}

bool upload_batch_state(World& w)
{
	if (w.subCount == 0)
		return true;
	hipStream_t s = w.stream;
	int32_t n = (int32_t)w.bodySubHost.size();
	if (w.bodySubDirty && n > 0)
	{
		const int32_t S = w.subCount;
		if (!w.bodySub.ensure(n) || !w.subBodyStartD.ensure((size_t)S + 1))
			return false;
		if (!hip_ok(hipMemcpyAsync(w.bodySub.p, w.bodySubHost.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s),
					"sub-world upload") ||
			!hip_ok(hipMemcpyAsync(w.subBodyStartD.p, w.subBodyStart.data(), ((size_t)S + 1) * sizeof(int32_t),
								   hipMemcpyHostToDevice, s),
					"sub-world upload"))
			return false;
		w.maxSubBodies = 0;
		for (int32_t k = 0; k < S; ++k)
			w.maxSubBodies = std::max(w.maxSubBodies, w.subBodyStart[k + 1] - w.subBodyStart[k]);
		w.bodySubDirty = false;
	}
	if (w.anySubDirtyHost)
	{
		int32_t S = w.subCount;
		if (!hip_ok(hipMemcpyAsync(w.subDirtyUpload.p, w.subDirtyHost.data(), S, hipMemcpyHostToDevice, s),
					"sub-world dirt upload"))
			return false;
		k_sub_or_dirty<<<grid_for(S), kBlock, 0, s>>>(w.subStable.p, w.subDirtyUpload.p, S);
		if (!hip_ok(hipStreamSynchronize(s), "sub-world dirt upload"))
			return false;
		std::fill(w.subDirtyHost.begin(), w.subDirtyHost.end(), 0);
		w.anySubDirtyHost = false;
	}
	return true;
}

// rebuild = force || (dirty && steps since the last build >= 4); a rebuild
// clears dirty and force and restarts the count, else the count advances
__global__ void k_stable_plan(uint32_t* subStable, int32_t S, uint32_t* counters)
{
	int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= S)
		return;
	uint32_t u = subStable[t];
	uint32_t since = u >> 8;
	bool rebuild = (u & SUB_FORCE) || ((u & SUB_DIRTY) && since >= 4u);
	if (rebuild)
	{
		subStable[t] = SUB_REBUILD;
		atomicAdd(&counters[CT_SUB_REBUILDS], 1u);
	}
	else
	{
		since = since < (1u << 20) ? since + 1u : since;
		subStable[t] = (u & (SUB_DIRTY | SUB_FORCE)) | (since << 8);
	}
}

void launch_stable_plan(World& w)
{
	if (w.subCount > 0)
		k_stable_plan<<<grid_for(w.subCount), kBlock, 0, w.stream>>>(w.subStable.p, w.subCount, w.counters.p);
}

__global__ void k_keep_flags(const uint32_t* bodies, int64_t n, const int32_t* bodySub, const uint32_t* subStable,
							 uint32_t* keep)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		keep[i] = (subStable[bodySub[bodies[i]]] & SUB_REBUILD) ? 0u : 1u;
}

__global__ void k_compact_entries(const uint64_t* keys, const uint32_t* bodies, const uint32_t* keep,
								  const uint32_t* offs, int64_t n, uint64_t* keysOut, uint32_t* bodiesOut)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && keep[i])
	{
		keysOut[offs[i]] = keys[i];
		bodiesOut[offs[i]] = bodies[i];
	}
}

bool build_stable_batched(World& w)
{
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	int32_t S = w.subCount;
	DBodies B = body_view(w);
	int64_t oldE = w.stableCount;
	if (n == 0)
		return true;
	if (!w.u32a.ensure(n) || !w.u32b.ensure(n) || !w.u32c.ensure(n) || !w.u32d.ensure(n) || !w.i32a.ensure(n) ||
		!w.i32d.ensure(std::max<int64_t>(oldE, 1)) || !w.i32e.ensure(std::max<int64_t>(oldE, 1)) ||
		!ensure_prim_scratch(w, std::max<int64_t>(n, oldE)))
		return false;
	uint32_t* cnt = w.u32a.p;
	uint32_t* offs = w.u32b.p;
	uint32_t* ov = w.u32c.p;
	uint32_t* ovOff = w.u32d.p;
	uint32_t* keep = (uint32_t*)w.i32d.p;
	uint32_t* keepOff = (uint32_t*)w.i32e.p;
	k_tier_count<<<grid_for(n), kBlock, 0, s>>>(B, n, BP_STABLE, 1, w.def.gridCellSize, cnt, ov);
	scan_exclusive2(cnt, offs, w.counters.p + CT_ENTRIES, ov, ovOff, w.counters.p + CT_OVERSIZED, n, w.scanStatus.p, s);
	k_compact_indices<<<grid_for(n), kBlock, 0, s>>>(ov, ovOff, n, w.i32a.p);
	if (oldE > 0)
	{
		k_keep_flags<<<grid_for(oldE), kBlock, 0, s>>>(w.stableBodies.p, oldE, w.bodySub.p, w.subStable.p, keep);
		scan_exclusive(keep, keepOff, oldE, w.counters.p + CT_KEPT, w.scanStatus.p, s);
	}
	else
	{
		hipMemsetAsync(w.counters.p + CT_KEPT, 0, sizeof(uint32_t), s);
	}
	w.subStableHost.resize(S);
	hipMemcpyAsync(w.subStableHost.data(), w.subStable.p, S * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
	if (!sync_counters(w))
		return false;
	int64_t eNew = w.hostCounters[CT_ENTRIES];
	int32_t nov = (int32_t)w.hostCounters[CT_OVERSIZED];
	int64_t eKeep = w.hostCounters[CT_KEPT];
	int64_t E = eKeep + eNew;

	// oversized list: kept entries of the other sub-worlds, new ones of the
	// rebuilding sub-worlds, ascending by body
	std::vector<int32_t> merged;
	merged.reserve(w.stableOversized.size() + nov);
	for (int32_t b : w.stableOversized)
		if (!(w.subStableHost[w.bodySubHost[b]] & SUB_REBUILD))
			merged.push_back(b);
	if (nov > 0)
	{
		size_t k0 = merged.size();
		merged.resize(k0 + nov);
		if (!hip_ok(hipMemcpy(merged.data() + k0, w.i32a.p, nov * sizeof(int32_t), hipMemcpyDeviceToHost),
					"stable oversized"))
			return false;
	}
	std::sort(merged.begin(), merged.end());
	w.stableOversized = merged;
	if (!merged.empty())
	{
		if (!w.dStableOversized.ensure(merged.size()))
			return false;
		hipMemcpyAsync(w.dStableOversized.p, w.stableOversized.data(), merged.size() * sizeof(int32_t),
					   hipMemcpyHostToDevice, s);
	}
	k_mark_cells<<<grid_for(n), kBlock, 0, s>>>(B, n, BP_STABLE, 1, BP_IN_STABLE_CELLS, cnt);
	if (E == 0)
	{
		w.stableCount = 0;
		return hip_ok(hipStreamSynchronize(s), "stable tier");
	}
	if (!w.stableKeysNext.ensure(E) || !w.stableBodiesNext.ensure(E) || !w.sortKeysAlt.ensure(E) ||
		!w.sortValsAlt.ensure(E) || !ensure_prim_scratch(w, E))
		return false;
	k_bounds_init<<<1, 64, 0, s>>>(w.counters.p);
	if (eKeep > 0)
	{
		k_compact_entries<<<grid_for(oldE), kBlock, 0, s>>>(w.stableKeys.p, w.stableBodies.p, keep, keepOff, oldE,
															  w.stableKeysNext.p, w.stableBodiesNext.p);
		k_key_bounds<<<reduce_grid(eKeep, kBlock), kBlock, 0, s>>>(w.stableKeysNext.p, eKeep, 1, w.counters.p);
	}
	if (eNew > 0)
		k_tier_emit<<<reduce_grid(n, kBlock), kBlock, 0, s>>>(B, n, BP_STABLE, 1, w.def.gridCellSize, cnt, offs,
												   w.stableKeysNext.p + eKeep, w.stableBodiesNext.p + eKeep,
												   eNew, w.counters.p, nullptr);
	if (!sync_counters(w))
		return false;
	int bits = 0;
	KeyCompress kc = key_compress_from(w.hostCounters, S, &bits);
	// kept entries precede new ones, and each sub-world's entries all come
	// from one of the two, so the stable sort keeps (key, body) order
	k_key_compress<<<grid_for(E), kBlock, 0, s>>>(w.stableKeysNext.p, E, kc);
	radix_sort_pairs(w.stableKeysNext.p, w.stableBodiesNext.p, w.sortKeysAlt.p, w.sortValsAlt.p, E, bits,
					 w.primScratch.p, w.scanStatus.p, s);
	k_key_expand<<<grid_for(E), kBlock, 0, s>>>(w.stableKeysNext.p, E, kc);
	std::swap(w.stableKeys, w.stableKeysNext);
	std::swap(w.stableBodies, w.stableBodiesNext);
	w.stableCount = E;
	return true;
}

__global__ void k_new_sub(const int32_t* bodySub, const uint64_t* keys, int32_t K, uint64_t* subKey, uint32_t* idx)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k < K)
	{
		subKey[k] = (uint64_t)(uint32_t)bodySub[(int32_t)(keys[k] >> 32)];
		idx[k] = (uint32_t)k;
	}
}

__global__ void k_gather_keys(const uint64_t* keys, const uint32_t* idx, int32_t K, uint64_t* out)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k < K)
		out[k] = keys[idx[k]];
}

bool stage_new_contacts(World& w, const uint64_t* keys, int32_t K)
{
	hipStream_t s = w.stream;
	w.newCount = K;
	if (K == 0)
		return true;
	if (!w.newKeys.ensure(K) || !w.s64a.ensure(K) || !w.s64b.ensure(K) || !w.s32a.ensure(K) || !w.s32b.ensure(K) ||
		!ensure_prim_scratch(w, K))
		return false;
	k_new_sub<<<grid_for(K), kBlock, 0, s>>>(w.bodySub.p, keys, K, w.s64a.p, w.s32a.p);
	radix_sort_pairs(w.s64a.p, w.s32a.p, w.s64b.p, w.s32b.p, K, bits_for((uint64_t)w.subCount), w.primScratch.p,
					 w.scanStatus.p, s);
	k_gather_keys<<<grid_for(K), kBlock, 0, s>>>(keys, w.s32a.p, K, w.newKeys.p);
	return true;
}

__device__ __forceinline__ int32_t contact_sub(const int32_t* bodySub, const int2* bodies, int32_t c)
{
	return bodySub[bodies[c].x];
}

// Removed contacts per sub-world (a sub-world's contacts are one range, so a
// wave's lanes share one or two keys) and in all; runs on reduce_grid
// workgroups, every lane taking the same number of turns.
__global__ void k_count_removed(const uint32_t* rm, const int2* bodies, const int32_t* bodySub, int32_t C,
								uint32_t* removed, uint32_t* counters)
{
	uint32_t n = 0;
	for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < C; base += (int64_t)gridDim.x * blockDim.x)
	{
		int32_t c = (int32_t)(base + threadIdx.x);
		bool r = c < C && rm[c] != 0u;
		wave_add_keyed(removed, r ? (uint32_t)contact_sub(bodySub, bodies, c) : ~0u, 1u);
		n += r ? 1u : 0u;
	}
	block_add<1>(counters, { CT_REMOVED }, { n });
}

__global__ void k_count_new(const uint64_t* keys, int32_t K, const int32_t* bodySub, uint32_t* added)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	wave_add_keyed(added, k < K ? (uint32_t)bodySub[(int32_t)(keys[k] >> 32)] : ~0u, 1u);
}

// contacts per sub-world after the step's removals and additions (entry S
// stays zero, so exclusive scans over S + 1 entries end with the total)
__global__ void k_sub_sizes(const int32_t* start, const uint32_t* removed, const uint32_t* added, int32_t S,
							uint32_t* newC)
{
	int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (t < S)
		newC[t] = (uint32_t)(start[t + 1] - start[t]) + added[t] - removed[t];
	else if (t == S)
		newC[t] = 0u;
}

// removal bookkeeping per sub-world range: a hole is a removed contact that
// stays inside the shrunk range, a tail a survivor beyond it
__global__ void k_hole_tail_sub(const uint32_t* rm, const int2* bodies, const int32_t* bodySub, const int32_t* start,
								const uint32_t* newC, int32_t C, uint32_t* hole, uint32_t* tail)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	int32_t sw = contact_sub(bodySub, bodies, c);
	uint32_t l = (uint32_t)(c - start[sw]);
	bool r = rm[c] != 0;
	hole[c] = (r && l < newC[sw]) ? 1u : 0u;
	tail[c] = (!r && l >= newC[sw]) ? 1u : 0u;
}

struct PlaceCtx
{
	const int32_t* bodySub;
	const int2* bodies;      // old contacts
	const int32_t* start;    // old ranges (S + 1)
	const uint32_t* newC;
	const uint32_t* newStart; // S + 1
	const uint32_t* newKeyStart;
	const uint32_t* rm;
	const uint32_t* holeOff;
	const uint32_t* tailOff;
	const int32_t* holeList;  // removed positions inside the shrunk ranges, ascending
	const uint32_t* counters;
	const uint32_t* regionOff; // per sub-world: start of its tail region in rankOf (its removals)
	const int32_t* rankOf;     // per tail-region slot: rank of its element among the tails
	int32_t C0;
};

// The swap-with-last removal of src/world.cpp:1722-1734 on one sub-world's
// range: the reference removes in descending order, moving the current last
// contact into each removed slot. The removed slots at or beyond the shrunk
// size newC come first and may take in a contact from the end too, so the
// survivors of the tail region [newC, oldC + added) (old slots, then the
// contacts created this step) stand in the order a replay of that region
// leaves them, and the k-th hole (ascending) takes the k-th of them. The
// region has as many slots as the sub-world has removals; one lane per
// sub-world replays it and records each slot's rank (-1: removed).
__global__ void k_tail_replay(const int32_t* start, const uint32_t* newC, const uint32_t* added, const uint32_t* rm,
							  const uint32_t* regionOff, int32_t S, int32_t* order, int32_t* rankOf)
{
	int32_t sw = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (sw >= S)
		return;
	const int32_t st0 = start[sw];
	const int32_t c0 = start[sw + 1] - st0;
	const int32_t nc = (int32_t)newC[sw];
	const int32_t n = c0 + (int32_t)added[sw] - nc;
	int32_t* ord = order + regionOff[sw];
	int32_t* rank = rankOf + regionOff[sw];
	for (int32_t i = 0; i < n; ++i)
	{
		ord[i] = i;
		rank[i] = -1;
	}
	int32_t top = n - 1;
	for (int32_t r = n - 1; r >= 0; --r)
	{
		int32_t l = nc + r;
		if (l < c0 && rm[st0 + l] != 0u)
		{
			if (r != top)
				ord[r] = ord[top];
			--top;
		}
	}
	for (int32_t k = 0; k <= top; ++k)
		rank[ord[k]] = k;
}

// Destination of the element at tail-region slot r of sub-world sw.
__device__ __forceinline__ int32_t place_tail(const PlaceCtx& X, int32_t sw, uint32_t r)
{
	int32_t st0 = X.start[sw];
	int32_t t = X.rankOf[X.regionOff[sw] + r];
	int32_t h = X.holeList[X.holeOff[st0] + t];
	return (int32_t)X.newStart[sw] + (h - st0);
}

// One lane per (surviving contact, point).
__global__ void k_place_old(PlaceCtx X, DContacts src, DContacts dstC)
{
	int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	int32_t c = (int32_t)(t >> 2);
	if (c >= X.C0 || X.rm[c])
		return;
	int32_t sw = contact_sub(X.bodySub, X.bodies, c);
	int32_t st0 = X.start[sw];
	uint32_t l = (uint32_t)(c - st0);
	int32_t d;
	if (l < X.newC[sw])
		d = (int32_t)X.newStart[sw] + (int32_t)l;
	else
		d = place_tail(X, sw, l - X.newC[sw]);
	// point slots at or beyond the count are never read (a regeneration
	// writes its points before any stage reads them): not copied
	const int j = (int)(t & 3);
	copy_contact_part(src, c, dstC, d, j, j < f2i(src.normal[c].w));
}

__global__ void k_place_new(PlaceCtx X, const uint64_t* keys, int32_t K, int32_t* dst)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= K)
		return;
	int32_t sw = X.bodySub[(int32_t)(keys[k] >> 32)];
	int32_t st0 = X.start[sw], st1 = X.start[sw + 1];
	uint32_t c0 = (uint32_t)(st1 - st0);
	uint32_t l = c0 + ((uint32_t)k - X.newKeyStart[sw]);
	uint32_t nc = X.newC[sw];
	if (l < nc)
	{
		dst[k] = (int32_t)X.newStart[sw] + (int32_t)l;
		return;
	}
	dst[k] = place_tail(X, sw, l - nc);
}

// Touching contacts a batch step removes (end events), and the gather of
// them: (body a, body b, sensor, sub-world of body a).
__global__ void k_end_flags(const uint32_t* rm, const float4* mat, int32_t C, uint32_t* ev)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	ev[c] = (rm[c] && (f2u(mat[c].w) & CF_TOUCHING)) ? 1u : 0u;
}

__global__ void k_gather_events_sub(const int2* bodies, const float4* mat, const int32_t* bodySub, const int32_t* idx,
									const uint32_t* count, int4* out)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= (int32_t)*count)
		return;
	int32_t c = idx[k];
	int2 b = bodies[c];
	uint32_t cf = f2u(mat[c].w);
	out[k] = make_int4(b.x, b.y, (cf & CF_SENSOR) ? 1 : 0, bodySub[b.x]);
}

// Contacts of one body (a destroy in a batch).
__global__ void k_body_contact_flags(const int2* bodies, int32_t C, int32_t body, uint32_t* flags)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	int2 ab = bodies[c];
	flags[c] = (ab.x == body || ab.y == body) ? 1u : 0u;
}

static bool remove_and_place_impl(World& w, int32_t destroyed, std::vector<int4>* events);

bool remove_and_place(World& w) { return remove_and_place_impl(w, -1, nullptr); }

// m3d_body_destroy's contact removal in a batch (src/world.cpp:576-593 in
// the body's sub-world): the body's contacts leave its sub-world's run in
// descending position, each swapped with the run's current last contact;
// the touching ones come back as (body a, body b, sensor, sub-world) in that
// order.
bool batch_destroy_contacts(World& w, int32_t body, std::vector<int4>& events)
{
	events.clear();
	launch_fill_u32(w.counters.p + CT_REMOVED, 0u, 1, w.stream);
	return remove_and_place_impl(w, body, &events);
}

// The batch step's removals (or, with destroyed >= 0, the removal of that
// body's contacts) and placement of the new contacts, per sub-world run.
static bool remove_and_place_impl(World& w, int32_t destroyed, std::vector<int4>* events)
{
	hipStream_t s = w.stream;
	int32_t S = w.subCount;
	int32_t C0 = w.contactCount;
	int32_t K = destroyed >= 0 ? 0 : w.newCount;
	if (C0 == 0 && K == 0)
		return true;
	if (!w.subWork.ensure(5 * (size_t)S + 9) || !w.u32a.ensure(std::max(C0, 1)) || !w.u32b.ensure(std::max(C0, 1)) ||
		!w.u32c.ensure(std::max(C0, 1)) || !w.i32a.ensure(std::max(C0, 1)) || !w.i32b.ensure(std::max(C0, 1)) ||
		!w.i32c.ensure(std::max(K, 1)) || !ensure_prim_scratch(w, std::max<int64_t>(C0, S + 1)))
		return false;
	uint32_t* removed = (uint32_t*)w.subWork.p;
	uint32_t* added = removed + S;
	uint32_t* newC = added + S;            // S + 1
	uint32_t* newStart = newC + S + 1;     // S + 1
	uint32_t* regionOff = newStart + S + 1; // S + 1
	if (!w.i32d.ensure((size_t)S + 1))
		return false;
	uint32_t* newKeyStart = (uint32_t*)w.i32d.p; // S + 1
	launch_fill_u32(removed, 0u, 2 * (int64_t)S, s);
	const int2* oldBodies = w.cc.bodies.p;
	uint32_t* rm = w.u32a.p;
	if (C0 > 0)
	{
		if (destroyed >= 0)
			k_body_contact_flags<<<grid_for(C0), kBlock, 0, s>>>(oldBodies, C0, destroyed, rm);
		else
			k_remove_flags<<<grid_for(C0), kBlock, 0, s>>>(body_view(w), oldBodies, C0, rm);
		k_count_removed<<<reduce_grid(C0, kBlock), kBlock, 0, s>>>(rm, oldBodies, w.bodySub.p, C0, removed,
																  w.counters.p);
	}
	if (K > 0)
		k_count_new<<<grid_for(K), kBlock, 0, s>>>(w.newKeys.p, K, w.bodySub.p, added);
	if (!sync_counters(w))
		return false;
	if (w.hostCounters[CT_REMOVED] == 0 && K == 0)
		return true;
	// end events of the touching removed contacts (at most R), back with the
	// placement's last wait
	int4* evHost = nullptr;
	if (w.hostCounters[CT_REMOVED] > 0)
	{
		const int32_t Rm = (int32_t)w.hostCounters[CT_REMOVED];
		if (!w.bEvFlag.ensure(C0) || !w.bEvOff.ensure(C0) || !w.bEvIdx.ensure(Rm) || !w.bEvOut.ensure(Rm) ||
			!ensure_pinned(w.pinnedRemoval, w.pinnedRemovalCap, (size_t)Rm * sizeof(int4), "end events"))
			return false;
		evHost = (int4*)w.pinnedRemoval;
		k_end_flags<<<grid_for(C0), kBlock, 0, s>>>(rm, w.cc.mat.p, C0, w.bEvFlag.p);
		scan_exclusive(w.bEvFlag.p, w.bEvOff.p, C0, w.counters.p + CT_END_EVENTS, w.scanStatus.p, s);
		k_compact_indices<<<grid_for(C0), kBlock, 0, s>>>(w.bEvFlag.p, w.bEvOff.p, C0, w.bEvIdx.p);
		k_gather_events_sub<<<grid_for(Rm), kBlock, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, w.bodySub.p, w.bEvIdx.p,
															w.counters.p + CT_END_EVENTS, w.bEvOut.p);
		if (!hip_ok(hipMemcpyAsync(evHost, w.bEvOut.p, (size_t)Rm * sizeof(int4), hipMemcpyDeviceToHost, s),
					"end events"))
			return false;
	}
	k_sub_sizes<<<grid_for(S + 1), kBlock, 0, s>>>(w.subContactStart.p, removed, added, S, newC);
	scan_exclusive(newC, newStart, S + 1, w.counters.p + CT_PLACED, w.scanStatus.p, s);
	// new keys are grouped by sub-world: exclusive scan of the counts
	// (added has S entries; the scan's total lands in the last slot)
	scan_exclusive(added, newKeyStart, S, newKeyStart + S, w.scanStatus.p, s);
	// tail regions, one slot per removal of the sub-world
	const int32_t R = (int32_t)w.hostCounters[CT_REMOVED];
	scan_exclusive(removed, regionOff, S, regionOff + S, w.scanStatus.p, s);
	if (!w.tailOrder.ensure(2 * (size_t)std::max(R, 1)))
		return false;
	k_tail_replay<<<grid_for(S), kBlock, 0, s>>>(w.subContactStart.p, newC, added, rm, regionOff, S, w.tailOrder.p,
												 w.tailOrder.p + std::max(R, 1));
	uint32_t* hole = w.u32b.p;
	uint32_t* tail = w.u32c.p;
	uint32_t* holeOff = (uint32_t*)w.i32a.p;
	uint32_t* tailOff = (uint32_t*)w.i32b.p;
	int32_t H = 0;
	if (C0 > 0)
	{
		k_hole_tail_sub<<<grid_for(C0), kBlock, 0, s>>>(rm, oldBodies, w.bodySub.p, w.subContactStart.p, newC, C0,
														hole, tail);
		scan_exclusive(hole, holeOff, C0, w.counters.p + CT_HOLES, w.scanStatus.p, s);
		scan_exclusive(tail, tailOff, C0, w.counters.p + CT_TAILS, w.scanStatus.p, s);
	}
	if (!sync_counters(w))
		return false;
	int32_t total = (int32_t)w.hostCounters[CT_PLACED];
	H = C0 > 0 ? (int32_t)w.hostCounters[CT_HOLES] : 0;
	if (H > 0)
	{
		if (!w.eventIdx.ensure(H))
			return false;
		k_compact_indices<<<grid_for(C0), kBlock, 0, s>>>(hole, holeOff, C0, w.eventIdx.p);
	}
	if (!ensure_contact_columns(w, w.cc2, w.contactCap2, std::max(total, 1), 0))
		return false;
	PlaceCtx X;
	X.bodySub = w.bodySub.p;
	X.bodies = oldBodies;
	X.start = w.subContactStart.p;
	X.newC = newC;
	X.newStart = newStart;
	X.newKeyStart = newKeyStart;
	X.rm = rm;
	X.holeOff = holeOff;
	X.tailOff = tailOff;
	X.holeList = w.eventIdx.p;
	X.counters = w.counters.p;
	X.regionOff = regionOff;
	X.rankOf = w.tailOrder.p + std::max(R, 1);
	X.C0 = C0;
	DContacts src = contact_view(w);
	DContacts dstC = columns_view(w.cc2, w.contactCap2);
	if (C0 > 0)
		k_place_old<<<grid_for(4 * (int64_t)C0), kBlock, 0, s>>>(X, src, dstC);
	if (K > 0)
	{
		int32_t* dst = w.i32c.p;
		k_place_new<<<grid_for(K), kBlock, 0, s>>>(X, w.newKeys.p, K, dst);
		k_create<<<grid_for(K), kBlock, 0, s>>>(body_view(w), dstC, w.newKeys.p, K, 0, dst);
	}
	hipMemcpyAsync(w.subContactStart.p, newStart, ((size_t)S + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
	std::swap(w.cc, w.cc2);
	std::swap(w.contactCap, w.contactCap2);
	w.contactCount = total;
	if (!hip_ok(hipStreamSynchronize(s), "contact placement"))
		return false;
	// each sub-world's end events in its own world's order: its removals in
	// descending contact position (src/world.cpp:1722-1734); sub-worlds in
	// ascending order (their contacts are ascending runs)
	if (evHost != nullptr)
	{
		const int32_t NE = (int32_t)w.hostCounters[CT_END_EVENTS];
		for (int32_t hi = 0; hi < NE;)
		{
			int32_t lo = hi;
			while (hi < NE && evHost[hi].w == evHost[lo].w)
				++hi;
			for (int32_t k = hi - 1; k >= lo; --k)
			{
				const int4& e = evHost[k];
				if (events != nullptr)
				{
					events->push_back(e);
					continue;
				}
				m3d_contact_event ce;
				ce.bodyA = { e.x, w.bodies[e.x].generation };
				ce.bodyB = { e.y, w.bodies[e.y].generation };
				ce.isSensor = e.z != 0;
				w.endEvents.push_back(ce);
			}
		}
	}
	return true;
}

} // namespace mx
