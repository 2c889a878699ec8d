// Shared definitions of the MI355X step backend: math (host+device), engine
// constants, SoA device layouts and the per-thread working structs.
//
// The public header's inline math is compiled for both host and device so
// that kernels evaluate exactly the same floating-point expressions, in the
// same order, as the reference engine (include/melon3d.h:82-233). All
// translation units are built with -ffp-contract=off and correctly rounded
// division/sqrt, so every +,-,*,/,sqrt matches the strict-IEEE reference
// build bit for bit; only sinf/cosf/atan2f may differ in the last ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#if defined(__HIP__)
#pragma clang force_cuda_host_device begin
#include "melon3d.h"
#pragma clang force_cuda_host_device end
#else
#include "melon3d.h" // host-only translation units (the API layer)
#endif

#define MX_HD __host__ __device__ __forceinline__

namespace mx
{

// ------------------------------------------------------------ constants
// Values from src/m3d_internal.h:31-63 and the local constexprs cited where used.
constexpr float kLinearSlop = 0.005f;
constexpr float kSpeculativeBase = 4.0f * kLinearSlop;
constexpr float kMaxDepenetrationSpeed = 3.0f;
constexpr float kFrictionRelaxation = 0.5f;
constexpr float kAABBMargin = 0.1f;
constexpr float kTimeToSleep = 0.5f;
constexpr float kSleepLinearVelSq = 0.05f * 0.05f;
constexpr float kSleepAngularVelSq = 0.05f * 0.05f;
constexpr float kSleepEnergyThresh = kSleepLinearVelSq + kSleepAngularVelSq;
constexpr float kSleepEnergyAlpha = 0.25f;
constexpr float kSleepWakeFactor = 3.0f;
constexpr int32_t kOversizedCellSpan = 64;
constexpr int32_t kBigIslandContacts = 300;
constexpr int32_t kMaxGraphColors = 8;

// body flag word (device) -----------------------------------------------
constexpr uint32_t BF_TYPE_MASK = 3u;       // m3d_body_type
constexpr uint32_t BF_SHAPE_SHIFT = 2;      // m3d_shape_type in bits 2..3
constexpr uint32_t BF_SENSOR = 1u << 4;
constexpr uint32_t BF_ENABLE_SLEEP = 1u << 5;
constexpr uint32_t BF_AWAKE = 1u << 6;
constexpr uint32_t BF_IN_USE = 1u << 7;

MX_HD int32_t bf_type(uint32_t f) { return (int32_t)(f & BF_TYPE_MASK); }
MX_HD int32_t bf_shape(uint32_t f) { return (int32_t)((f >> BF_SHAPE_SHIFT) & 3u); }

// reference sidecar bits (src/m3d_internal.h:271-274)
constexpr uint8_t SB_DYNAMIC = 1;
constexpr uint8_t SB_AWAKE = 2;
constexpr uint8_t SB_IN_USE = 4;
constexpr uint8_t SB_KINEMATIC = 8;

// broadphase per-body bits
constexpr uint8_t BP_HAS_FAT = 1;
constexpr uint8_t BP_IN_STABLE_TIER = 2;
constexpr uint8_t BP_ACTIVE = 4;
constexpr uint8_t BP_STABLE = 8;
constexpr uint8_t BP_MOVED = 16;
constexpr uint8_t BP_IN_ACTIVE_CELLS = 32;
constexpr uint8_t BP_IN_STABLE_CELLS = 64;

// contact flag bits
constexpr uint32_t CF_TOUCHING = 1u;
constexpr uint32_t CF_SENSOR = 2u;
constexpr uint32_t CF_TUNNEL_GUARD = 4u;
constexpr uint32_t CF_CURVED = 8u;
constexpr uint32_t CF_HAS_CACHE = 16u;

// ----------------------------------------------------------- small math
MX_HD float v3_get(m3d_vec3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
MX_HD void v3_set(m3d_vec3& v, int i, float s)
{
	if (i == 0)
		v.x = s;
	else if (i == 1)
		v.y = s;
	else
		v.z = s;
}
MX_HD m3d_vec3 f4_v3(float4 f) { return m3d_v3(f.x, f.y, f.z); }
MX_HD m3d_quat f4_q(float4 f)
{
	m3d_quat q = { f.x, f.y, f.z, f.w };
	return q;
}
MX_HD float4 v3_f4(m3d_vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
MX_HD float4 q_f4(m3d_quat q) { return make_float4(q.x, q.y, q.z, q.w); }
MX_HD float i2f(int32_t i) { return __builtin_bit_cast(float, i); }
MX_HD int32_t f2i(float f) { return __builtin_bit_cast(int32_t, f); }
MX_HD float u2f(uint32_t i) { return __builtin_bit_cast(float, i); }
MX_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }

MX_HD uint64_t contact_key(int32_t a, int32_t b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; }

// Grid cell of a coordinate and the packed cell key (src/grid.cpp:19-37):
// floor(v / cell) with an offset of 2^20, 21 bits per axis, x above y above z.
MX_HD int64_t cell_coord(float v, float inv) { return (int64_t)floorf(v * inv); }

MX_HD uint64_t pack_cell(int64_t x, int64_t y, int64_t z)
{
	const int64_t off = (int64_t)1 << 20;
	uint64_t ux = (uint64_t)(x + off) & 0x1FFFFFu;
	uint64_t uy = (uint64_t)(y + off) & 0x1FFFFFu;
	uint64_t uz = (uint64_t)(z + off) & 0x1FFFFFu;
	return (ux << 42) | (uy << 21) | uz;
}

// ------------------------------------------------------------- shapes
struct Shape
{
	int32_t type;
	float radius;
	float halfHeight;
	m3d_vec3 halfExtents;
};

// --------------------------------------------------------- manifolds
// Working copy of a contact manifold (src/m3d_internal.h:154-194).
struct MPoint
{
	m3d_vec3 localAnchorA;
	m3d_vec3 localAnchorB;
	m3d_vec3 frictionAnchorA;
	m3d_vec3 frictionAnchorB;
	float separation;
	uint32_t id;
	float lambdaN;
	float lambdaT;
	float separationOffset;
	m3d_vec3 slipDir;
};

// A manifold under construction: point j's fields are lane j of 4-wide
// vector registers, so a point taken or stored at a run-time index stays in
// registers (an array of MPoint would go to per-lane scratch memory).
typedef float mx_float4v __attribute__((ext_vector_type(4)));
typedef uint32_t mx_uint4v __attribute__((ext_vector_type(4)));
struct Manifold
{
	m3d_vec3 normal;
	int32_t pointCount;
	mx_float4v ax, ay, az, bx, by, bz, fax, fay, faz, fbx, fby, fbz, sep, lamN, lamT, off, slx, sly, slz;
	mx_uint4v id;
	MX_HD MPoint point(int j) const
	{
		MPoint p;
		p.localAnchorA = m3d_v3(ax[j], ay[j], az[j]);
		p.localAnchorB = m3d_v3(bx[j], by[j], bz[j]);
		p.frictionAnchorA = m3d_v3(fax[j], fay[j], faz[j]);
		p.frictionAnchorB = m3d_v3(fbx[j], fby[j], fbz[j]);
		p.separation = sep[j];
		p.id = id[j];
		p.lambdaN = lamN[j];
		p.lambdaT = lamT[j];
		p.separationOffset = off[j];
		p.slipDir = m3d_v3(slx[j], sly[j], slz[j]);
		return p;
	}
	MX_HD void set_point(int j, const MPoint& p)
	{
		ax[j] = p.localAnchorA.x;
		ay[j] = p.localAnchorA.y;
		az[j] = p.localAnchorA.z;
		bx[j] = p.localAnchorB.x;
		by[j] = p.localAnchorB.y;
		bz[j] = p.localAnchorB.z;
		fax[j] = p.frictionAnchorA.x;
		fay[j] = p.frictionAnchorA.y;
		faz[j] = p.frictionAnchorA.z;
		fbx[j] = p.frictionAnchorB.x;
		fby[j] = p.frictionAnchorB.y;
		fbz[j] = p.frictionAnchorB.z;
		sep[j] = p.separation;
		id[j] = p.id;
		lamN[j] = p.lambdaN;
		lamT[j] = p.lambdaT;
		off[j] = p.separationOffset;
		slx[j] = p.slipDir.x;
		sly[j] = p.slipDir.y;
		slz[j] = p.slipDir.z;
	}
};

// ----------------------------------------------------- device layouts
// Bodies: structure of float4 columns indexed by body slot.
struct DBodies
{
	float4* pos;        // position xyz, invMass
	float4* rot;        // rotation
	float4* vel;        // linear velocity xyz, gravityScale
	float4* angVel;     // angular velocity xyz, mass
	float4* prevPos;    // previous substep position xyz
	float4* prevRot;
	float4* prevVel;    // previous substep linear velocity
	float4* prevAngVel; // previous substep angular velocity
	float4* invIL;      // inverse local inertia xyz, linearDamping
	float4* iw0;        // inverse world inertia column x, angularDamping
	float4* iw1;        // column y, sleepTime
	float4* iw2;        // column z, smoothedEnergy
	float4* shape;      // halfExtents xyz, radius
	float4* mat;        // halfHeight, friction, restitution, rollingResistance
	float4* force;      // force xyz
	float4* torque;     // torque xyz
	uint32_t* flags;    // BF_* bits
	uint4* filter;      // categoryBits, maskBits, groupIndex, -
	float4* fatLo;      // fat AABB lower xyz
	float4* fatHi;      // fat AABB upper xyz
	uint8_t* bp;        // BP_* bits
	uint8_t* sb;        // reference sidecar bits (SB_*)
	// batched worlds (null for a plain world): sub-world of each body, and
	// per sub-world stable-tier state (SUB_* bits, steps since the last
	// stable build from bit 8)
	const int32_t* sub;
	uint32_t* subStable;
};

constexpr uint32_t SUB_DIRTY = 1u;   // stable tier membership changed
constexpr uint32_t SUB_FORCE = 2u;   // stable tier must be rebuilt now
constexpr uint32_t SUB_REBUILD = 4u; // stable tier is rebuilt this step

#if defined(__HIP__)
// A sub-world's stable tier is dirty (the per-world stableDirty of
// src/world.cpp:2106, 130, 565 when worlds are batched).
__device__ __forceinline__ void mark_sub_stable_dirty(const DBodies& B, int32_t i)
{
	if (B.sub != nullptr)
		atomicOr(&B.subStable[B.sub[i]], SUB_DIRTY);
}
#endif

// Contacts: contact columns indexed by contact slot; manifold point columns
// indexed [pointSlot * cap + contact] so a wave reading point j of 64
// consecutive contacts is one coalesced access.
struct DContacts
{
	int32_t cap;
	// point j of contact c is at c * cs + j * js of the point columns: the
	// world's columns keep a contact's four points together (cs 4, js 1), so
	// a contact gathered at random costs one 64-byte run per column
	int32_t cs, js;
	int2* bodies;     // bodyA < bodyB
	float4* mat;      // friction, restitution, rollingResistance, flags (CF_* bits)
	float4* normal;   // normal xyz, pointCount (int bits)
	float4* cache;    // pose cache, 4 per contact (see cache_of)
	float4* pA;  // localAnchorA xyz, separationOffset
	float4* pB;  // localAnchorB xyz, wN
	float4* pJA; // angJA xyz, lambdaN
	float4* pJB; // angJB xyz, lambdaT
	float4* pRA; // rA xyz, separation
	float4* pRB; // rB xyz, id (bits)
	float4* pFA; // frictionAnchorA xyz
	float4* pFB; // frictionAnchorB xyz
	float4* pSlip; // slipDir xyz
	int2* hint;    // schedule warm start: last color, last overflow level
};

MX_HD Shape load_shape(const DBodies& B, int32_t i)
{
	float4 s = B.shape[i];
	float4 m = B.mat[i];
	Shape sh;
	sh.type = bf_shape(B.flags[i]);
	sh.radius = s.w;
	sh.halfHeight = m.x;
	sh.halfExtents = m3d_v3(s.x, s.y, s.z);
	return sh;
}

MX_HD m3d_transform load_xf(const DBodies& B, int32_t i)
{
	m3d_transform xf;
	xf.p = f4_v3(B.pos[i]);
	xf.q = f4_q(B.rot[i]);
	return xf;
}

MX_HD m3d_mat3 load_iw(const DBodies& B, int32_t i)
{
	m3d_mat3 m;
	m.cx = f4_v3(B.iw0[i]);
	m.cy = f4_v3(B.iw1[i]);
	m.cz = f4_v3(B.iw2[i]);
	return m;
}

MX_HD void store_iw(const DBodies& B, int32_t i, const m3d_mat3& m)
{
	B.iw0[i] = v3_f4(m.cx, B.iw0[i].w);
	B.iw1[i] = v3_f4(m.cy, B.iw1[i].w);
	B.iw2[i] = v3_f4(m.cz, B.iw2[i].w);
}

MX_HD int32_t pidx(const DContacts& C, int32_t c, int32_t j) { return c * C.cs + j * C.js; }

// Pose cache of contact c as one 64-byte run: position A, rotation A,
// position B, rotation B.
MX_HD float4* cache_of(const DContacts& C, int32_t c) { return C.cache + 4 * (int64_t)c; }

// Load the persistent manifold of contact c (normal, count, points).
MX_HD void load_manifold(const DContacts& C, int32_t c, Manifold& m)
{
	float4 n = C.normal[c];
	m.normal = f4_v3(n);
	m.pointCount = f2i(n.w);
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = pidx(C, c, j);
		float4 a = C.pA[p], b = C.pB[p], ja = C.pJA[p], jb = C.pJB[p], ra = C.pRA[p], rb = C.pRB[p];
		float4 fa = C.pFA[p], fb = C.pFB[p], sl = C.pSlip[p];
		MPoint mp;
		mp.localAnchorA = f4_v3(a);
		mp.separationOffset = a.w;
		mp.localAnchorB = f4_v3(b);
		mp.lambdaN = ja.w;
		mp.lambdaT = jb.w;
		mp.separation = ra.w;
		mp.id = f2u(rb.w);
		mp.frictionAnchorA = f4_v3(fa);
		mp.frictionAnchorB = f4_v3(fb);
		mp.slipDir = f4_v3(sl);
		m.set_point(j, mp);
	}
}

// The stored manifold of contact c read where it lies: normal and count at
// once, a point's fields when point(j) is called (no local copy of the four
// points, which a kernel would have to keep in scratch memory).
struct StoredManifold
{
	const DContacts* C;
	int32_t c;
	m3d_vec3 normal;
	int32_t pointCount;
	MX_HD StoredManifold(const DContacts& cs, int32_t ci) : C(&cs), c(ci)
	{
		float4 n = cs.normal[ci];
		normal = f4_v3(n);
		pointCount = f2i(n.w);
	}
	MX_HD MPoint point(int j) const
	{
		int32_t p = pidx(*C, c, j);
		float4 a = C->pA[p], b = C->pB[p], ja = C->pJA[p], jb = C->pJB[p], ra = C->pRA[p], rb = C->pRB[p];
		float4 fa = C->pFA[p], fb = C->pFB[p], sl = C->pSlip[p];
		MPoint mp;
		mp.localAnchorA = f4_v3(a);
		mp.separationOffset = a.w;
		mp.localAnchorB = f4_v3(b);
		mp.lambdaN = ja.w;
		mp.lambdaT = jb.w;
		mp.separation = ra.w;
		mp.id = f2u(rb.w);
		mp.frictionAnchorA = f4_v3(fa);
		mp.frictionAnchorB = f4_v3(fb);
		mp.slipDir = f4_v3(sl);
		return mp;
	}
};

// Store the persistent fields of a manifold's live points. Slots at or
// beyond pointCount keep their previous contents, as the reference's
// in-place manifold regeneration does (src/collide.cpp:813-836); no stage
// reads them.
MX_HD void store_manifold(const DContacts& C, int32_t c, const Manifold& m)
{
	C.normal[c] = v3_f4(m.normal, i2f(m.pointCount));
	for (int j = 0; j < m.pointCount; ++j)
	{
		int32_t p = pidx(C, c, j);
		const MPoint mp = m.point(j);
		C.pA[p] = v3_f4(mp.localAnchorA, mp.separationOffset);
		float4 b = C.pB[p];
		C.pB[p] = make_float4(mp.localAnchorB.x, mp.localAnchorB.y, mp.localAnchorB.z, b.w);
		float4 ja = C.pJA[p];
		C.pJA[p] = make_float4(ja.x, ja.y, ja.z, mp.lambdaN);
		float4 jb = C.pJB[p];
		C.pJB[p] = make_float4(jb.x, jb.y, jb.z, mp.lambdaT);
		float4 ra = C.pRA[p];
		C.pRA[p] = make_float4(ra.x, ra.y, ra.z, mp.separation);
		float4 rb = C.pRB[p];
		C.pRB[p] = make_float4(rb.x, rb.y, rb.z, u2f(mp.id));
		C.pFA[p] = v3_f4(mp.frictionAnchorA, 0.0f);
		C.pFB[p] = v3_f4(mp.frictionAnchorB, 0.0f);
		C.pSlip[p] = v3_f4(mp.slipDir, 0.0f);
	}
}

} // namespace mx
