// Scene queries on the device: batched ray casts over the world's grid
// tiers, one lane per ray, each the reference's RayCastInternal
// (src/world.cpp:1081-1200): the coarse lists first (active, then stable),
// then the cells along the segment by the Amanatides & Woo walk, each
// cell's entries found by binary search in the sorted active tier, then the
// stable tier; a body is taken when its hit fraction is strictly below the
// best so far. The reference skips a body it has already tested in this
// query; testing it again cannot change the result (same fraction, strict
// comparison), so the lanes keep no per-query marks.
#include "mx_steps.h"

#include <float.h>
#include <mutex>
#include <vector>

namespace mx
{

DBodies make_body_view(World& w);

struct RayCtx
{
	DBodies B;
	const uint64_t* aKeys;
	const uint32_t* aBodies;
	int64_t aCount;
	const uint64_t* sKeys;
	const uint32_t* sBodies;
	int64_t sCount;
	const int32_t* aOver;
	int32_t nAOver;
	const int32_t* sOver;
	int32_t nSOver;
	float cellSize;
	// bodies in use the grid may not find (created or moved since the tiers
	// were built, or whose tight box has left the fat box the tiers hold),
	// tested by every query that needs every overlapping body; null for the
	// ray cast, which follows the grid as the reference's does
	const int32_t* extra;
	const uint32_t* nExtra;
};

struct RayHit
{
	int32_t body;
	float fraction;
	m3d_vec3 normal;
};

__device__ __forceinline__ void ray_test(const RayCtx& R, int32_t bi, m3d_vec3 o, m3d_vec3 d, uint32_t mask,
										 float& maxFraction, RayHit& h)
{
	uint32_t fl = R.B.flags[bi];
	if (!(fl & BF_IN_USE) || (R.B.filter[bi].x & mask) == 0u)
		return;
	Shape s = load_shape(R.B, bi);
	m3d_transform xf = load_xf(R.B, bi);
	float f;
	m3d_vec3 n;
	if (ray_cast_shape(s, xf, o, d, &f, &n) && f < maxFraction)
	{
		maxFraction = f;
		h.body = bi;
		h.fraction = f;
		h.normal = n;
	}
}

// first index with keys[i] >= key
__device__ __forceinline__ int64_t lower_bound_key(const uint64_t* keys, int64_t n, uint64_t key)
{
	int64_t lo = 0, hi = n;
	while (lo < hi)
	{
		int64_t mid = (lo + hi) >> 1;
		if (keys[mid] < key)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

__device__ __forceinline__ void visit_cell(const RayCtx& R, int64_t x, int64_t y, int64_t z, m3d_vec3 o, m3d_vec3 d,
										   uint32_t mask, float& maxFraction, RayHit& h)
{
	float cs = R.cellSize;
	m3d_vec3 probe = m3d_v3(((float)x + 0.5f) * cs, ((float)y + 0.5f) * cs, ((float)z + 0.5f) * cs);
	float inv = 1.0f / cs;
	uint64_t key = pack_cell((int64_t)floorf(probe.x * inv), (int64_t)floorf(probe.y * inv),
							 (int64_t)floorf(probe.z * inv));
	for (int64_t i = lower_bound_key(R.aKeys, R.aCount, key); i < R.aCount && R.aKeys[i] == key; ++i)
		ray_test(R, (int32_t)R.aBodies[i], o, d, mask, maxFraction, h);
	for (int64_t i = lower_bound_key(R.sKeys, R.sCount, key); i < R.sCount && R.sKeys[i] == key; ++i)
		ray_test(R, (int32_t)R.sBodies[i], o, d, mask, maxFraction, h);
}

// rayO: origin xyz, mask bits (w); rayD: translation xyz. outA: point xyz,
// fraction (2 when nothing is hit); outB: normal xyz, body index bits (-1).
__global__ void k_ray_cast(RayCtx R, const float4* rayO, const float4* rayD, int32_t count, float4* outA, float4* outB)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= count)
		return;
	float4 O = rayO[k], D = rayD[k];
	m3d_vec3 o = f4_v3(O), d = f4_v3(D);
	uint32_t mask = f2u(O.w);
	RayHit h;
	h.body = -1;
	h.fraction = 2.0f;
	h.normal = m3d_v3_zero();
	float maxFraction = 1.0f;
	for (int32_t i = 0; i < R.nAOver; ++i)
		ray_test(R, R.aOver[i], o, d, mask, maxFraction, h);
	for (int32_t i = 0; i < R.nSOver; ++i)
		ray_test(R, R.sOver[i], o, d, mask, maxFraction, h);
	if (R.aCount > 0 || R.sCount > 0)
	{
		float cs = R.cellSize;
		float inv = 1.0f / cs;
		int64_t cell[3] = { (int64_t)floorf(o.x * inv), (int64_t)floorf(o.y * inv), (int64_t)floorf(o.z * inv) };
		const float dv[3] = { d.x, d.y, d.z };
		const float ov[3] = { o.x, o.y, o.z };
		int64_t stepDir[3];
		float tMax[3], tDelta[3];
#pragma unroll
		for (int a = 0; a < 3; ++a)
		{
			if (fabsf(dv[a]) < 1e-12f)
			{
				stepDir[a] = 0;
				tMax[a] = FLT_MAX;
				tDelta[a] = FLT_MAX;
			}
			else
			{
				stepDir[a] = dv[a] > 0.0f ? 1 : -1;
				float boundary = (float)(cell[a] + (stepDir[a] > 0 ? 1 : 0)) * cs;
				tMax[a] = (boundary - ov[a]) / dv[a];
				tDelta[a] = cs / fabsf(dv[a]);
			}
		}
		float t = 0.0f;
		for (int guard = 0; guard < 4096 && t <= maxFraction; ++guard)
		{
			visit_cell(R, cell[0], cell[1], cell[2], o, d, mask, maxFraction, h);
			int axis = 0;
			if (tMax[1] < tMax[axis])
				axis = 1;
			if (tMax[2] < tMax[axis])
				axis = 2;
			t = tMax[axis];
			if (t > 1.0f || stepDir[axis] == 0)
				break;
			cell[axis] += stepDir[axis];
			tMax[axis] += tDelta[axis];
		}
	}
	m3d_vec3 point = h.body >= 0 ? m3d_mul_add(o, h.fraction, d) : m3d_add(o, d);
	outA[k] = v3_f4(point, h.fraction);
	outB[k] = v3_f4(h.normal, i2f(h.body));
}

// One body's part of a sphere cast (src/world.cpp:1286-1392) evaluated on
// its own: kind 1 the swept sphere touches the body at the start, 2 a hit at
// *fraction, 0 none. The reference walks bodies in index order, returns on
// the first start overlap and otherwise keeps a hit only when its fraction
// is strictly below the best so far (its box loop also stops once past the
// best, which can only drop hits that would not be kept), so its result is:
// the lowest-index start overlap, else the smallest fraction with ties to
// the lowest index. Evaluating each body alone and applying that rule gives
// the same result in any candidate order.
struct CastHit
{
	int kind;
	float fraction;
	m3d_vec3 point, normal;
};

MX_HD CastHit sphere_cast_body(const Shape& sh, m3d_transform xf, m3d_vec3 origin, float radius, m3d_vec3 tr, float len)
{
	CastHit h;
	h.kind = 0;
	h.fraction = 2.0f;
	h.point = m3d_v3_zero();
	h.normal = m3d_v3_zero();
	{
		m3d_vec3 closest;
		float raw = point_shape_closest(sh, xf, origin, &closest);
		if (raw - radius < 1e-4f)
		{
			h.kind = 1;
			h.fraction = 0.0f;
			h.point = closest;
			m3d_vec3 outward = raw >= 0.0f ? m3d_sub(origin, closest) : m3d_sub(closest, origin);
			h.normal = m3d_normalize(outward);
			return h;
		}
	}
	if (sh.type == M3D_SHAPE_SPHERE || sh.type == M3D_SHAPE_CAPSULE)
	{
		Shape inflated = sh;
		inflated.radius += radius;
		float f;
		m3d_vec3 n;
		if (ray_cast_shape(inflated, xf, origin, tr, &f, &n))
		{
			m3d_vec3 center = m3d_mul_add(origin, f, tr);
			h.kind = 2;
			h.fraction = f;
			h.normal = n;
			h.point = m3d_mul_add(center, -radius, n);
		}
		return h;
	}
	float t = 0.0f;
	for (int it = 0; it < 128; ++it)
	{
		m3d_vec3 p = m3d_mul_add(origin, t, tr);
		m3d_vec3 closest;
		float raw = point_shape_closest(sh, xf, p, &closest);
		float d = raw - radius;
		if (d < 1e-4f)
		{
			h.kind = 2;
			h.fraction = t;
			h.point = closest;
			m3d_vec3 outward = raw >= 0.0f ? m3d_sub(p, closest) : m3d_sub(closest, p);
			h.normal = m3d_normalize(outward);
			return h;
		}
		if (len < 1e-9f)
			break;
		t += d / len;
		if (t > 1.0f)
			break;
	}
	return h;
}

struct CastBest
{
	int32_t startBody, hitBody;
	CastHit start, hit;
};

__device__ __forceinline__ void cast_test(const RayCtx& R, int32_t bi, m3d_vec3 o, float radius, m3d_vec3 d, float len,
										  uint32_t mask, const m3d_aabb& sweep, CastBest& B)
{
	uint32_t fl = R.B.flags[bi];
	if (!(fl & BF_IN_USE) || (R.B.filter[bi].x & mask) == 0u)
		return;
	if (B.startBody >= 0 && bi >= B.startBody)
		return; // a lower-index body already touches at the start
	Shape sh = load_shape(R.B, bi);
	m3d_transform xf = load_xf(R.B, bi);
	if (!m3d_aabb_overlap(sweep, shape_aabb(sh, xf)))
		return;
	CastHit h = sphere_cast_body(sh, xf, o, radius, d, len);
	if (h.kind == 1)
	{
		B.startBody = bi;
		B.start = h;
	}
	else if (h.kind == 2 && (B.hitBody < 0 || h.fraction < B.hit.fraction ||
							 (h.fraction == B.hit.fraction && bi < B.hitBody)))
	{
		B.hitBody = bi;
		B.hit = h;
	}
}

// Sphere casts: candidates are the coarse lists, the bodies the grid may not
// find (R.extra) and the bodies of the grid cells the swept sphere's box
// covers (any other body whose tight box meets the sweep has an entry in one
// of them: its tight box lies in its fat box); sweeps over more than
// kCastCells cells test every body instead. castO: origin xyz, mask bits; castD: translation
// xyz, radius. Outputs as k_ray_cast.
constexpr int64_t kCastCells = 512;

__global__ void k_sphere_cast(RayCtx R, int32_t nBodies, const float4* castO, const float4* castD, int32_t count,
							  float4* outA, float4* outB)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= count)
		return;
	float4 O = castO[k], D = castD[k];
	m3d_vec3 o = f4_v3(O), d = f4_v3(D);
	uint32_t mask = f2u(O.w);
	float radius = D.w > 0.0f ? D.w : 0.0f;
	m3d_vec3 end = m3d_add(o, d);
	m3d_vec3 pad = m3d_v3(radius + 1e-3f, radius + 1e-3f, radius + 1e-3f);
	m3d_aabb sweep;
	sweep.lowerBound = m3d_sub(m3d_min(o, end), pad);
	sweep.upperBound = m3d_add(m3d_max(o, end), pad);
	float len = m3d_length(d);
	CastBest B;
	B.startBody = -1;
	B.hitBody = -1;
	float inv = 1.0f / R.cellSize;
	int64_t x0 = cell_coord(sweep.lowerBound.x, inv), x1 = cell_coord(sweep.upperBound.x, inv);
	int64_t y0 = cell_coord(sweep.lowerBound.y, inv), y1 = cell_coord(sweep.upperBound.y, inv);
	int64_t z0 = cell_coord(sweep.lowerBound.z, inv), z1 = cell_coord(sweep.upperBound.z, inv);
	int64_t cells = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
	if (cells > kCastCells || cells <= 0)
	{
		for (int32_t bi = 0; bi < nBodies; ++bi)
			cast_test(R, bi, o, radius, d, len, mask, sweep, B);
	}
	else
	{
		for (int32_t i = 0; i < R.nAOver; ++i)
			cast_test(R, R.aOver[i], o, radius, d, len, mask, sweep, B);
		for (int32_t i = 0; i < R.nSOver; ++i)
			cast_test(R, R.sOver[i], o, radius, d, len, mask, sweep, B);
		if (R.extra != nullptr)
			for (uint32_t i = 0; i < *R.nExtra; ++i)
				cast_test(R, R.extra[i], o, radius, d, len, mask, sweep, B);
		for (int64_t x = x0; x <= x1; ++x)
			for (int64_t y = y0; y <= y1; ++y)
				for (int64_t z = z0; z <= z1; ++z)
				{
					uint64_t key = pack_cell(x, y, z);
					for (int64_t i = lower_bound_key(R.aKeys, R.aCount, key); i < R.aCount && R.aKeys[i] == key; ++i)
						cast_test(R, (int32_t)R.aBodies[i], o, radius, d, len, mask, sweep, B);
					for (int64_t i = lower_bound_key(R.sKeys, R.sCount, key); i < R.sCount && R.sKeys[i] == key; ++i)
						cast_test(R, (int32_t)R.sBodies[i], o, radius, d, len, mask, sweep, B);
				}
	}
	int32_t body = B.startBody >= 0 ? B.startBody : B.hitBody;
	const CastHit* h = B.startBody >= 0 ? &B.start : &B.hit;
	if (body >= 0)
	{
		outA[k] = v3_f4(h->point, h->fraction);
		outB[k] = v3_f4(h->normal, i2f(body));
	}
	else
	{
		outA[k] = v3_f4(end, 2.0f);
		outB[k] = v3_f4(m3d_v3_zero(), i2f(-1));
	}
}

// ---- AABB overlap queries ----------------------------------------------

// flag[i] = body i is in use and the grid may not find it: created after the
// tiers were built (slot at or beyond gridN, or no fat box), or its tight
// box is not inside the fat box the tiers hold (moved by the API, or by the
// step past its fat box).
__global__ void k_query_extra_flags(DBodies B, int32_t n, int32_t gridN, uint32_t* flag)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	uint32_t f = 0u;
	if (B.flags[i] & BF_IN_USE)
	{
		m3d_aabb t = shape_aabb(load_shape(B, i), load_xf(B, i));
		float4 lo = B.fatLo[i], hi = B.fatHi[i];
		bool inside = i < gridN && (B.bp[i] & BP_HAS_FAT) && lo.x <= t.lowerBound.x && lo.y <= t.lowerBound.y &&
					  lo.z <= t.lowerBound.z && t.upperBound.x <= hi.x && t.upperBound.y <= hi.y &&
					  t.upperBound.z <= hi.z;
		f = inside ? 0u : 1u;
	}
	flag[i] = f;
}

__global__ void k_query_extra_compact(const uint32_t* flag, const uint32_t* offs, int32_t n, int32_t* extra)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n && flag[i])
		extra[offs[i]] = i;
}

// The bodies whose tight box overlaps query k's box
// (src/world.cpp:1394-1416), from the coarse lists, R.extra and the grid
// cells the box covers (every body, for boxes over more than kCastCells
// cells or brute != 0). A body can be found more than once (cells, tiers,
// lists); the caller sorts and drops repeats. Pass 0 counts into cnt, pass
// 1 writes (query << 32 | body) at offs[k].
__device__ __forceinline__ void overlap_emit(const RayCtx& R, int32_t bi, const m3d_aabb& q, uint64_t k, int pass,
											 uint32_t& n, uint64_t* out, uint32_t base)
{
	if (!(R.B.flags[bi] & BF_IN_USE))
		return;
	if (!m3d_aabb_overlap(q, shape_aabb(load_shape(R.B, bi), load_xf(R.B, bi))))
		return;
	if (pass)
		out[base + n] = (k << 32) | (uint32_t)bi;
	++n;
}

__global__ void k_overlap(RayCtx R, int32_t nBodies, int brute, const float4* qLo, const float4* qHi, int32_t count,
						  int pass, uint32_t* cnt, const uint32_t* offs, uint64_t* out)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= count)
		return;
	m3d_aabb q;
	q.lowerBound = f4_v3(qLo[k]);
	q.upperBound = f4_v3(qHi[k]);
	uint32_t n = 0u;
	uint32_t base = pass ? offs[k] : 0u;
	float inv = 1.0f / R.cellSize;
	int64_t x0 = cell_coord(q.lowerBound.x, inv), x1 = cell_coord(q.upperBound.x, inv);
	int64_t y0 = cell_coord(q.lowerBound.y, inv), y1 = cell_coord(q.upperBound.y, inv);
	int64_t z0 = cell_coord(q.lowerBound.z, inv), z1 = cell_coord(q.upperBound.z, inv);
	int64_t cells = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
	if (brute || cells > kCastCells || cells <= 0)
	{
		for (int32_t bi = 0; bi < nBodies; ++bi)
			overlap_emit(R, bi, q, (uint64_t)k, pass, n, out, base);
	}
	else
	{
		for (int32_t i = 0; i < R.nAOver; ++i)
			overlap_emit(R, R.aOver[i], q, (uint64_t)k, pass, n, out, base);
		for (int32_t i = 0; i < R.nSOver; ++i)
			overlap_emit(R, R.sOver[i], q, (uint64_t)k, pass, n, out, base);
		for (uint32_t i = 0; i < *R.nExtra; ++i)
			overlap_emit(R, R.extra[i], q, (uint64_t)k, pass, n, out, base);
		for (int64_t x = x0; x <= x1; ++x)
			for (int64_t y = y0; y <= y1; ++y)
				for (int64_t z = z0; z <= z1; ++z)
				{
					uint64_t key = pack_cell(x, y, z);
					for (int64_t i = lower_bound_key(R.aKeys, R.aCount, key); i < R.aCount && R.aKeys[i] == key; ++i)
						overlap_emit(R, (int32_t)R.aBodies[i], q, (uint64_t)k, pass, n, out, base);
					for (int64_t i = lower_bound_key(R.sKeys, R.sCount, key); i < R.sCount && R.sKeys[i] == key; ++i)
						overlap_emit(R, (int32_t)R.sBodies[i], q, (uint64_t)k, pass, n, out, base);
				}
	}
	if (!pass)
		cnt[k] = n;
}

__global__ void k_first_of_run(const uint64_t* keys, int64_t n, uint32_t* flag)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_compact_keys(const uint64_t* keys, const uint32_t* flag, const uint32_t* offs, int64_t n,
							   uint64_t* out)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flag[i])
		out[offs[i]] = keys[i];
}

} // namespace mx

using namespace mx;

// ---- body transforms of one frame ----------------------------------------

// out[k] = pose of body idx[k] (7 floats: position, rotation xyzw); idx -1:
// the identity pose at the origin.
__global__ void k_gather_poses(DBodies B, const int32_t* idx, int32_t count, float* out)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= count)
		return;
	int32_t i = idx[k];
	float4 p = i >= 0 ? B.pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	float4 q = i >= 0 ? B.rot[i] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
	float* o = out + 7 * (int64_t)k;
	o[0] = p.x;
	o[1] = p.y;
	o[2] = p.z;
	o[3] = q.x;
	o[4] = q.y;
	o[5] = q.z;
	o[6] = q.w;
}

// ---- allocation-free ray casts ------------------------------------------

// Staging of m3dx_world_ray_cast_batch_noalloc, one array per GPU for the
// life of the process: rays k of a call use float4 slots k (origin, mask),
// kNoAllocMaxRays + k (translation), 2 kNoAllocMaxRays + k and
// 3 kNoAllocMaxRays + k (the two result words). Every slot a call reads is
// written by the same call first (its rays' inputs by the host copies, its
// results by the kernel), so nothing needs clearing between calls; calls
// that overlap on one GPU share the array and race.
constexpr int32_t kNoAllocMaxRays = 65536; // 4 MiB per GPU
static __device__ __align__(256) float4 scratch[4 * kNoAllocMaxRays];

// Device address of `scratch` on a device, looked up once per device with
// that device current (the array's host-side name is not a device address).
static float4* scratch_on(int device)
{
	constexpr int kMaxDevices = 64;
	static float4* addr[kMaxDevices];
	static std::mutex lock;
	if (device < 0 || device >= kMaxDevices)
	{
		set_error("ray cast batch (no allocation): device %d out of range", device);
		return nullptr;
	}
	std::lock_guard<std::mutex> g(lock);
	if (addr[device] == nullptr)
	{
		int prev = 0;
		if (!hip_ok(hipGetDevice(&prev), "hipGetDevice") || !hip_ok(hipSetDevice(device), "hipSetDevice"))
			return nullptr;
		void* p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		(void)hipSetDevice(prev);
		if (!hip_ok(e, "hipGetSymbolAddress(scratch)") || p == nullptr)
			return nullptr;
		addr[device] = (float4*)p;
	}
	return addr[device];
}

static RayCtx ray_ctx(World& w)
{
	RayCtx R;
	R.B = make_body_view(w);
	R.aKeys = w.activeKeys.p;
	R.aBodies = w.activeBodies.p;
	R.aCount = w.activeCount;
	R.sKeys = w.stableKeys.p;
	R.sBodies = w.stableBodies.p;
	R.sCount = w.stableCount;
	R.aOver = w.dActiveOversized.p;
	R.nAOver = (int32_t)w.activeOversized.size();
	R.sOver = w.dStableOversized.p;
	R.nSOver = (int32_t)w.stableOversized.size();
	R.cellSize = w.def.gridCellSize;
	R.extra = nullptr;
	R.nExtra = nullptr;
	return R;
}

// The bodies the grid may not find (RayCtx::extra), listed on the device; the
// count stays on the device (*nExtra).
static bool query_extra(World& w, RayCtx& R)
{
	int32_t n = w.deviceBodyCount;
	hipStream_t s = w.stream;
	if (!w.qFlag.ensure((size_t)std::max(n, 1)) || !w.qOffs.ensure((size_t)n + 1) ||
		!w.qExtra.ensure((size_t)std::max(n, 1)) || !ensure_prim_scratch(w, std::max(n, 1)))
		return false;
	uint32_t* total = w.qOffs.p + n;
	if (n == 0)
	{
		launch_fill_u32(total, 0u, 1, s);
	}
	else
	{
		k_query_extra_flags<<<(n + 255) / 256, 256, 0, s>>>(make_body_view(w), n, w.gridBodies, w.qFlag.p);
		scan_exclusive(w.qFlag.p, w.qOffs.p, n, total, w.scanStatus.p, s);
		k_query_extra_compact<<<(n + 255) / 256, 256, 0, s>>>(w.qFlag.p, w.qOffs.p, n, w.qExtra.p);
	}
	R.extra = w.qExtra.p;
	R.nExtra = total;
	return true;
}

// host side of the batched queries: arrays up, kernel, results down
template <class Launch>
static int run_query(World& w, const std::vector<float4>& o, const std::vector<float4>& d, m3d_ray_result* out,
					 bool withExtra, Launch&& launch)
{
	int32_t count = (int32_t)o.size();
	if (w.gridDirty || w.stableDirty)
	{
		if (!push_host_changes(w) || !refresh_broadphase(w, 0.0f))
			return -1;
		w.hostStale = true;
	}
	hipStream_t s = w.stream;
	DBuf<float4> buf;
	if (!buf.ensure(4 * (size_t)count))
		return -1;
	float4* dO = buf.p;
	float4* dD = dO + count;
	float4* dA = dD + count;
	float4* dB = dA + count;
	RayCtx R = ray_ctx(w);
	if (withExtra && !query_extra(w, R))
		return -1;
	std::vector<float4> a((size_t)count), b((size_t)count);
	bool ok = hip_ok(hipMemcpyAsync(dO, o.data(), count * sizeof(float4), hipMemcpyHostToDevice, s), "queries") &&
			  hip_ok(hipMemcpyAsync(dD, d.data(), count * sizeof(float4), hipMemcpyHostToDevice, s), "queries");
	if (ok)
	{
		launch(R, dO, dD, count, dA, dB, s);
		ok = hip_ok(hipGetLastError(), "query") &&
			 hip_ok(hipMemcpyAsync(a.data(), dA, count * sizeof(float4), hipMemcpyDeviceToHost, s), "hits") &&
			 hip_ok(hipMemcpyAsync(b.data(), dB, count * sizeof(float4), hipMemcpyDeviceToHost, s), "hits") &&
			 hip_ok(hipStreamSynchronize(s), "query");
	}
	buf.release();
	if (!ok)
		return -1;
	for (int32_t k = 0; k < count; ++k)
	{
		m3d_ray_result& r = out[k];
		int32_t bi = f2i(b[k].w);
		r.hit = bi >= 0;
		r.body = bi >= 0 ? m3d_body_id{ bi, w.bodies[bi].generation } : m3d_body_id_null();
		r.point = m3d_v3(a[k].x, a[k].y, a[k].z);
		r.fraction = a[k].w;
		r.normal = m3d_v3(b[k].x, b[k].y, b[k].z);
	}
	return 0;
}

namespace mx
{

// Sorted, repeat-free (query << 32 | body) pairs of the bodies whose tight
// box overlaps each query box (src/world.cpp:1394-1416 per query), on the
// device: host changes are pushed first (no broadphase refresh: the
// reference's overlap query reads no grid), candidates come from the grid
// plus the bodies it may not find, or from every body for a batch (whose
// cell keys carry the sub-world) and for large boxes.
bool overlap_query(World& w, const m3d_aabb* boxes, int32_t count, std::vector<uint64_t>& pairs)
{
	pairs.clear();
	if (count <= 0)
		return true;
	if (!push_host_changes(w))
		return false;
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	RayCtx R = ray_ctx(w);
	if (!query_extra(w, R))
		return false;
	std::vector<float4> lo((size_t)count), hi((size_t)count);
	for (int32_t k = 0; k < count; ++k)
	{
		lo[k] = v3_f4(boxes[k].lowerBound, 0.0f);
		hi[k] = v3_f4(boxes[k].upperBound, 0.0f);
	}
	DBuf<char> buf;
	size_t qb = (size_t)count * sizeof(float4);
	size_t cb = ((size_t)count + 1) * sizeof(uint32_t);
	if (!buf.ensure(2 * qb + 2 * cb + 512))
		return false;
	float4* dLo = (float4*)buf.p;
	float4* dHi = dLo + count;
	uint32_t* cnt = (uint32_t*)(dHi + count);
	uint32_t* offs = cnt + count + 1;
	const int brute = w.subCount > 0 ? 1 : 0;
	const unsigned grid = (unsigned)((count + 127) / 128);
	bool ok = hip_ok(hipMemcpyAsync(dLo, lo.data(), qb, hipMemcpyHostToDevice, s), "overlap boxes") &&
			  hip_ok(hipMemcpyAsync(dHi, hi.data(), qb, hipMemcpyHostToDevice, s), "overlap boxes");
	if (ok)
	{
		k_overlap<<<grid, 128, 0, s>>>(R, n, brute, dLo, dHi, count, 0, cnt, nullptr, nullptr);
		scan_exclusive(cnt, offs, count, w.counters.p + CT_RESPAWN_REMOVED, w.scanStatus.p, s);
		ok = hip_ok(hipGetLastError(), "overlap count") && sync_counters(w);
	}
	int64_t total = ok ? (int64_t)w.hostCounters[CT_RESPAWN_REMOVED] : 0;
	if (ok && total > 0)
	{
		int bitsQ = 1;
		while ((1ll << bitsQ) < (int64_t)count)
			++bitsQ;
		// (growing a buffer frees the old one: none of these is one the
		// launches below read through R, which holds qExtra and qOffs)
		ok = w.qPairs.ensure(total) && w.qPairsAlt.ensure(total) && w.qVals.ensure(total) &&
			 w.qValsAlt.ensure(total) && w.qUFlag.ensure(total) && w.qUOffs.ensure(total + 1) &&
			 ensure_prim_scratch(w, total);
		if (ok)
		{
			k_overlap<<<grid, 128, 0, s>>>(R, n, brute, dLo, dHi, count, 1, cnt, offs, w.qPairs.p);
			launch_fill_u32(w.qVals.p, 0u, total, s);
			radix_sort_pairs(w.qPairs.p, w.qVals.p, w.qPairsAlt.p, w.qValsAlt.p, total, 32 + bitsQ, w.primScratch.p,
							 w.scanStatus.p, s);
			unsigned g2 = (unsigned)((total + 255) / 256);
			k_first_of_run<<<g2, 256, 0, s>>>(w.qPairs.p, total, w.qUFlag.p);
			scan_exclusive(w.qUFlag.p, w.qUOffs.p, total, w.counters.p + CT_RESPAWN_REMOVED, w.scanStatus.p, s);
			k_compact_keys<<<g2, 256, 0, s>>>(w.qPairs.p, w.qUFlag.p, w.qUOffs.p, total, w.qPairsAlt.p);
			ok = hip_ok(hipGetLastError(), "overlap pairs") && sync_counters(w);
		}
		if (ok)
		{
			int64_t uniq = (int64_t)w.hostCounters[CT_RESPAWN_REMOVED];
			pairs.resize((size_t)uniq);
			ok = hip_ok(hipMemcpy(pairs.data(), w.qPairsAlt.p, (size_t)uniq * sizeof(uint64_t), hipMemcpyDeviceToHost),
						"overlap pairs");
		}
	}
	buf.release();
	return ok;
}

} // namespace mx

extern "C"
{

// Many AABB overlap queries in one call on the device (no counterpart in the
// reference, which runs one query per call with a callback): the bodies of
// query k, in ascending body index as the reference's callback would see
// them (src/world.cpp:1394-1416), are out[offsets[k] .. offsets[k + 1]).
// offsets (count + 1 entries) is always filled; out receives at most cap ids
// (out may be null with cap 0). Returns the total number of results, or -1.
int m3dx_world_overlap_aabb_batch(m3d_world* world, const m3d_aabb* boxes, int32_t count, int32_t* offsets,
								  m3d_body_id* out, int32_t cap)
{
	if (world == nullptr || count < 0 || (count > 0 && (boxes == nullptr || offsets == nullptr)) || cap < 0 ||
		(cap > 0 && out == nullptr))
	{
		set_error("overlap batch: bad arguments");
		return -1;
	}
	World& w = world->w;
	std::vector<uint64_t> pairs;
	if (!overlap_query(w, boxes, count, pairs))
		return -1;
	if ((int64_t)pairs.size() > INT32_MAX)
	{
		set_error("overlap batch: more than 2^31 results");
		return -1;
	}
	size_t j = 0;
	for (int32_t k = 0; k < count; ++k)
	{
		offsets[k] = (int32_t)j;
		while (j < pairs.size() && (int32_t)(pairs[j] >> 32) == k)
		{
			if ((int64_t)j < cap)
			{
				int32_t bi = (int32_t)(uint32_t)pairs[j];
				out[j] = m3d_body_id{ bi, w.bodies[bi].generation };
			}
			++j;
		}
	}
	if (count >= 0)
		offsets[count] = (int32_t)j;
	return (int32_t)pairs.size();
}

// m3dx_world_ray_cast_batch without any allocation: device staging in the
// library's per-GPU `scratch` array (at most kNoAllocMaxRays rays per
// call), host staging in fixed chunks on the stack. The world's broadphase
// must be current (the grid as the last step left it, with no API change
// since), since refreshing it may grow device buffers. Same results as
// m3dx_world_ray_cast_batch. Returns 0 or -1.
int m3dx_world_ray_cast_batch_noalloc(m3d_world* world, const m3d_vec3* origins, const m3d_vec3* translations,
									  const uint32_t* maskBits, int32_t count, m3d_ray_result* out)
{
	if (world == nullptr || count < 0 || (count > 0 && (origins == nullptr || translations == nullptr || out == nullptr)))
	{
		set_error("ray cast batch (no allocation): null arrays");
		return -1;
	}
	if (count > kNoAllocMaxRays)
	{
		set_error("ray cast batch (no allocation): %d rays exceed the maximum of %d per call", count, kNoAllocMaxRays);
		return -1;
	}
	World& w = world->w;
	if (w.subCount > 0)
	{
		set_error("ray cast batch (no allocation): batched worlds are not supported");
		return -1;
	}
	// The reference refreshes the broadphase before a query while the grid
	// or the stable tier is dirty (src/world.cpp:1090-1093). That refresh is
	// not done here (it may grow device buffers). A dirty grid is refused.
	// A dirty stable tier alone is served when the grid's last refresh was a
	// query's (nothing moved since) and the stable tier is not due for a
	// rebuild: the refresh would then leave both tiers as they are and only
	// count one more refresh since the last stable build, which is done here.
	bool stableDue = w.stableForce || w.stepsSinceStableBuild >= 4;
	if (w.gridDirty || !w.dirtyList.empty() || (w.stableDirty && (!w.queryRefreshed || stableDue)))
	{
		set_error("ray cast batch (no allocation): the broadphase is not current (a query that may refresh it, such "
				  "as m3dx_world_ray_cast_batch, has to run first)");
		return -1;
	}
	if (count == 0)
		return 0;
	float4* sc = scratch_on(w.device);
	if (sc == nullptr)
		return -1;
	if (w.stableDirty)
		++w.stepsSinceStableBuild;
	float4* dO = sc;
	float4* dD = sc + kNoAllocMaxRays;
	float4* dA = sc + 2 * kNoAllocMaxRays;
	float4* dB = sc + 3 * kNoAllocMaxRays;
	hipStream_t s = w.stream;
	constexpr int32_t kChunk = 1024;
	float4 hO[kChunk], hD[kChunk];
	for (int32_t c0 = 0; c0 < count; c0 += kChunk)
	{
		int32_t m = std::min(kChunk, count - c0);
		for (int32_t k = 0; k < m; ++k)
		{
			const m3d_vec3 o = origins[c0 + k], d = translations[c0 + k];
			hO[k] = make_float4(o.x, o.y, o.z, u2f(maskBits != nullptr ? maskBits[c0 + k] : 0xFFFFFFFFu));
			hD[k] = make_float4(d.x, d.y, d.z, 0.0f);
		}
		// synchronous copies: the chunk buffers are reused by the next chunk
		if (!hip_ok(hipMemcpyAsync(dO + c0, hO, m * sizeof(float4), hipMemcpyHostToDevice, s), "rays") ||
			!hip_ok(hipMemcpyAsync(dD + c0, hD, m * sizeof(float4), hipMemcpyHostToDevice, s), "rays") ||
			!hip_ok(hipStreamSynchronize(s), "rays"))
			return -1;
	}
	k_ray_cast<<<(count + 127) / 128, 128, 0, s>>>(ray_ctx(w), dO, dD, count, dA, dB);
	if (!hip_ok(hipGetLastError(), "ray cast"))
		return -1;
	for (int32_t c0 = 0; c0 < count; c0 += kChunk)
	{
		int32_t m = std::min(kChunk, count - c0);
		if (!hip_ok(hipMemcpyAsync(hO, dA + c0, m * sizeof(float4), hipMemcpyDeviceToHost, s), "hits") ||
			!hip_ok(hipMemcpyAsync(hD, dB + c0, m * sizeof(float4), hipMemcpyDeviceToHost, s), "hits") ||
			!hip_ok(hipStreamSynchronize(s), "hits"))
			return -1;
		for (int32_t k = 0; k < m; ++k)
		{
			m3d_ray_result& r = out[c0 + k];
			int32_t bi = f2i(hD[k].w);
			r.hit = bi >= 0;
			r.body = bi >= 0 ? m3d_body_id{ bi, w.bodies[bi].generation } : m3d_body_id_null();
			r.point = m3d_v3(hO[k].x, hO[k].y, hO[k].z);
			r.fraction = hO[k].w;
			r.normal = m3d_v3(hD[k].x, hD[k].y, hD[k].z);
		}
	}
	return 0;
}

// Transforms of many bodies in one call: out[k] = m3d_body_transform(world,
// ids[k]) (include/melon3d.h:470, src/world.cpp:634-645), gathered on the
// device after a step instead of through a download of every body.
int m3dx_world_transforms(m3d_world* world, const m3d_body_id* ids, int32_t count, m3d_transform* out)
{
	if (world == nullptr || count < 0 || (count > 0 && (ids == nullptr || out == nullptr)))
	{
		set_error("transforms: null arrays");
		return -1;
	}
	World& w = world->w;
	if (count == 0)
		return 0;
	auto valid = [&](m3d_body_id id) {
		return id.index >= 0 && id.index < (int32_t)w.bodies.size() && w.bodies[id.index].inUse &&
			   w.bodies[id.index].generation == id.generation;
	};
	auto dirty = [&](int32_t i) { return i < (int32_t)w.dirtyFlag.size() && w.dirtyFlag[i] != 0; };
	if (!w.hostStale)
	{
		for (int32_t k = 0; k < count; ++k)
			out[k] = m3d_body_transform(world, ids[k]);
		return 0;
	}
	std::vector<int32_t> idx((size_t)count);
	for (int32_t k = 0; k < count; ++k)
	{
		int32_t i = ids[k].index;
		idx[k] = valid(ids[k]) && i < w.deviceBodyCount && !dirty(i) ? i : -1;
	}
	hipStream_t s = w.stream;
	DBuf<char> buf;
	if (!buf.ensure((size_t)count * (sizeof(int32_t) + 7 * sizeof(float)) + 256))
		return -1;
	int32_t* dIdx = (int32_t*)buf.p;
	float* dOut = (float*)(buf.p + (((size_t)count * sizeof(int32_t) + 255) & ~(size_t)255));
	std::vector<float> h((size_t)count * 7);
	bool ok = hip_ok(hipMemcpyAsync(dIdx, idx.data(), count * sizeof(int32_t), hipMemcpyHostToDevice, s), "poses");
	if (ok)
	{
		k_gather_poses<<<(count + 255) / 256, 256, 0, s>>>(make_body_view(w), dIdx, count, dOut);
		ok = hip_ok(hipGetLastError(), "poses") &&
			 hip_ok(hipMemcpyAsync(h.data(), dOut, (size_t)count * 7 * sizeof(float), hipMemcpyDeviceToHost, s),
					"poses") &&
			 hip_ok(hipStreamSynchronize(s), "poses");
	}
	buf.release();
	if (!ok)
		return -1;
	for (int32_t k = 0; k < count; ++k)
	{
		int32_t i = ids[k].index;
		if (idx[k] < 0 && valid(ids[k]))
		{
			// a record the host holds ahead of the device
			out[k].p = w.bodies[i].position;
			out[k].q = w.bodies[i].rotation;
			continue;
		}
		const float* o = h.data() + 7 * (size_t)k;
		out[k].p = m3d_v3(o[0], o[1], o[2]);
		out[k].q = { o[3], o[4], o[5], o[6] };
	}
	return 0;
}

// Batched ray casts on the device: out[k] is m3d_world_ray_cast_filtered(
// world, origins[k], translations[k], maskBits ? maskBits[k] : all) for
// every k, evaluated as the reference does (src/world.cpp:1081-1200) over
// the device grid. Plain worlds; a dirty broadphase is refreshed with
// dt = 0 first, as the reference does. Returns 0 or -1.
int m3dx_world_ray_cast_batch(m3d_world* world, const m3d_vec3* origins, const m3d_vec3* translations,
							  const uint32_t* maskBits, int32_t count, m3d_ray_result* out)
{
	if (world == nullptr || count < 0 || (count > 0 && (origins == nullptr || translations == nullptr || out == nullptr)))
	{
		set_error("ray cast batch: null arrays");
		return -1;
	}
	World& w = world->w;
	if (w.subCount > 0)
	{
		set_error("ray cast batch: batched worlds are not supported");
		return -1;
	}
	if (count == 0)
		return 0;
	std::vector<float4> o((size_t)count), d((size_t)count);
	for (int32_t k = 0; k < count; ++k)
	{
		uint32_t m = maskBits != nullptr ? maskBits[k] : 0xFFFFFFFFu;
		o[k] = make_float4(origins[k].x, origins[k].y, origins[k].z, u2f(m));
		d[k] = make_float4(translations[k].x, translations[k].y, translations[k].z, 0.0f);
	}
	return run_query(w, o, d, out, false,
					 [](const RayCtx& R, const float4* dO, const float4* dD, int32_t n, float4* dA, float4* dB,
						hipStream_t s) { k_ray_cast<<<(n + 127) / 128, 128, 0, s>>>(R, dO, dD, n, dA, dB); });
}

// Batched sphere casts on the device: out[k] = m3d_world_sphere_cast(world,
// origins[k], radii[k], translations[k], maskBits ? maskBits[k] : all),
// the reference's result (src/world.cpp:1286-1392) with candidates taken
// from the device grid. Plain worlds. Returns 0 or -1.
int m3dx_world_sphere_cast_batch(m3d_world* world, const m3d_vec3* origins, const float* radii,
								 const m3d_vec3* translations, const uint32_t* maskBits, int32_t count,
								 m3d_ray_result* out)
{
	if (world == nullptr || count < 0 ||
		(count > 0 && (origins == nullptr || radii == nullptr || translations == nullptr || out == nullptr)))
	{
		set_error("sphere cast batch: null arrays");
		return -1;
	}
	World& w = world->w;
	if (w.subCount > 0)
	{
		set_error("sphere cast batch: batched worlds are not supported");
		return -1;
	}
	if (count == 0)
		return 0;
	std::vector<float4> o((size_t)count), d((size_t)count);
	for (int32_t k = 0; k < count; ++k)
	{
		uint32_t m = maskBits != nullptr ? maskBits[k] : 0xFFFFFFFFu;
		o[k] = make_float4(origins[k].x, origins[k].y, origins[k].z, u2f(m));
		d[k] = make_float4(translations[k].x, translations[k].y, translations[k].z, radii[k]);
	}
	int32_t nBodies = w.deviceBodyCount;
	return run_query(w, o, d, out, true,
					 [nBodies](const RayCtx& R, const float4* dO, const float4* dD, int32_t n, float4* dA, float4* dB,
							   hipStream_t s) {
						 k_sphere_cast<<<(n + 127) / 128, 128, 0, s>>>(R, nBodies, dO, dD, n, dA, dB);
					 });
}

} // extern "C"
