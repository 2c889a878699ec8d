// Islands, island solve, sleeping and the end-of-step body passes.
//
//   islands   lock-free union-find (hooking always attaches the larger root
//             under the smaller, so every root is its component's smallest
//             body index) over touching non-sensor dynamic-dynamic contacts
//             and dynamic-dynamic joints; islands numbered in order of their
//             smallest body, as the reference's ascending scan assigns them
//             (src/world.cpp:1868-1958)
//   lists     island body/contact lists by a stable sort on the island id,
//             so each list stays ascending (src/world.cpp:1974-2010)
//   solve     one lane per awake island runs the reference island solve
//             (serial Gauss-Seidel, or greedy coloring with packet-form
//             normals and serial overflow for islands with >= 300 contacts)
//             (src/xpbd.cpp:898-974, :1167-1369)
//   sleep     low-pass energy, timers and whole-island sleep
//             (src/world.cpp:2060-2108)
#include "mx_steps.h"

#include <algorithm>

namespace mx
{

DBodies make_body_view(World& w);
DContacts make_contact_view(World& w);

constexpr int kBlockS = 256;
static inline unsigned grids(int64_t n) { return (unsigned)((n + kBlockS - 1) / kBlockS); }

// ----------------------------------------------------------- islands

__global__ void k_any_awake_dyn(const uint8_t* sb, int32_t n, uint32_t* counters)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n && (sb[i] & (SB_DYNAMIC | SB_AWAKE)) == (SB_DYNAMIC | SB_AWAKE))
		counters[CT_ANY_AWAKE_DYN] = 1u;
}

__global__ void k_uf_init(int32_t* parent, int32_t n)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n)
		parent[i] = i;
}

__device__ __forceinline__ int32_t uf_find(int32_t* parent, int32_t x)
{
	for (;;)
	{
		int32_t p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
		if (p == x)
			return x;
		int32_t gp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
		if (gp != p)
			__atomic_store_n(&parent[x], gp, __ATOMIC_RELAXED); // halving: still an ancestor
		x = p;
	}
}

__device__ __forceinline__ void uf_union(int32_t* parent, int32_t a, int32_t b)
{
	for (;;)
	{
		int32_t ra = uf_find(parent, a);
		int32_t rb = uf_find(parent, b);
		if (ra == rb)
			return;
		int32_t lo = ra < rb ? ra : rb;
		int32_t hi = ra < rb ? rb : ra;
		if (atomicCAS(&parent[hi], hi, lo) == hi)
			return;
	}
}

__global__ void k_uf_contacts(const int2* bodies, const float4* mat, const uint8_t* sb, int32_t C, int32_t* parent)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	uint32_t cf = f2u(mat[c].w);
	if (!(cf & CF_TOUCHING) || (cf & CF_SENSOR))
		return;
	int2 ab = bodies[c];
	if ((sb[ab.x] & SB_DYNAMIC) && (sb[ab.y] & SB_DYNAMIC))
		uf_union(parent, ab.x, ab.y);
}

__global__ void k_uf_joints(const DJoint* J, int32_t nJ, const uint32_t* flags, int32_t* parent)
{
	int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j >= nJ)
		return;
	const DJoint& jt = J[j];
	if (jt.inUse && bf_type(flags[jt.bodyA]) == M3D_BODY_DYNAMIC && bf_type(flags[jt.bodyB]) == M3D_BODY_DYNAMIC)
		uf_union(parent, jt.bodyA, jt.bodyB);
}

__global__ void k_uf_roots(int32_t* parent, const uint8_t* sb, int32_t n, uint32_t* rootFlag)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	int32_t r = uf_find(parent, i);
	parent[i] = r;
	rootFlag[i] = ((sb[i] & SB_DYNAMIC) && r == i) ? 1u : 0u;
}

__global__ void k_island_of(const int32_t* parent, const uint8_t* sb, const uint32_t* rootRank, int32_t n,
							int32_t* islandOf, uint32_t* islandAwake)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	uint8_t f = sb[i];
	if (!(f & SB_DYNAMIC))
	{
		islandOf[i] = -1;
		return;
	}
	int32_t isl = (int32_t)rootRank[parent[i]];
	islandOf[i] = isl;
	if (f & SB_AWAKE)
		islandAwake[isl] = 1u;
}

// WakeBodyInternal for every not-yet-awake body of an awake island
__global__ void k_island_wake(DBodies B, int32_t n, const int32_t* islandOf, const uint32_t* islandAwake,
							  uint32_t* counters)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	int32_t isl = islandOf[i];
	if (isl < 0 || !islandAwake[isl])
		return;
	if (B.flags[i] & BF_AWAKE)
		return;
	B.flags[i] |= BF_AWAKE;
	B.sb[i] |= SB_AWAKE;
	float4 a = B.iw1[i];
	a.w = 0.0f;
	B.iw1[i] = a;
	float4 b = B.iw2[i];
	b.w = 4.0f * kSleepEnergyThresh;
	B.iw2[i] = b;
	counters[CT_STABLE_DIRTY] = 1u;
}

// solver contacts: touching, non-sensor, dynamic body's island awake
__global__ void k_contact_island(const int2* bodies, const float4* mat, const float4* normal, const uint8_t* sb,
								 const int32_t* islandOf, const uint32_t* islandAwake, int32_t C, uint64_t* keys,
								 uint32_t* vals, uint32_t* flag, uint32_t* islandContacts, uint32_t* counters)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= C)
		return;
	uint32_t cf = f2u(mat[c].w);
	uint32_t f = 0;
	int32_t isl = -1;
	if (cf & CF_TOUCHING)
	{
		atomicAdd(&counters[CT_TOUCHING], 1u);
		if (!(cf & CF_SENSOR))
		{
			int2 ab = bodies[c];
			int32_t dyn = (sb[ab.x] & SB_DYNAMIC) ? ab.x : ab.y;
			isl = islandOf[dyn];
			if (isl >= 0 && islandAwake[isl])
			{
				f = 1;
				atomicAdd(&islandContacts[isl], 1u);
				atomicAdd(&counters[CT_SOLVED_CONTACTS], 1u);
				uint32_t pc = (uint32_t)f2i(normal[c].w);
				atomicAdd(&counters[CT_SOLVED_POINTS], pc);
				if (!(cf & CF_CURVED))
					atomicAdd(&counters[CT_BB_POINTS], pc);
			}
		}
	}
	flag[c] = f;
	keys[c] = f ? (uint64_t)isl : 0ull;
	vals[c] = (uint32_t)c;
}

__global__ void k_body_island_keys(const int32_t* islandOf, const uint32_t* islandAwake, int32_t n, uint64_t* keys,
								   uint32_t* vals, uint32_t* flag, uint32_t* islandBodies)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	int32_t isl = islandOf[i];
	uint32_t f = (isl >= 0 && islandAwake[isl]) ? 1u : 0u;
	flag[i] = f;
	keys[i] = f ? (uint64_t)isl : 0ull;
	vals[i] = (uint32_t)i;
	if (f)
		atomicAdd(&islandBodies[isl], 1u);
}

__global__ void k_compact_pairs(const uint64_t* keys, const uint32_t* vals, const uint32_t* flag, const uint32_t* offs,
								int64_t n, uint64_t* ko, uint32_t* vo)
{
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flag[i])
	{
		ko[offs[i]] = keys[i];
		vo[offs[i]] = vals[i];
	}
}

__global__ void k_joint_island(const DJoint* J, int32_t nJ, const uint32_t* flags, const int32_t* islandOf,
							   const uint32_t* islandAwake, int32_t* jointIsland)
{
	int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j >= nJ)
		return;
	const DJoint& jt = J[j];
	int32_t isl = -1;
	if (jt.inUse)
	{
		int32_t dyn = bf_type(flags[jt.bodyA]) == M3D_BODY_DYNAMIC ? jt.bodyA : jt.bodyB;
		if (bf_type(flags[dyn]) == M3D_BODY_DYNAMIC)
		{
			int32_t k = islandOf[dyn];
			if (k >= 0 && islandAwake[k])
				isl = k;
		}
	}
	jointIsland[j] = isl;
}

// ------------------------------------------------------- island solve

struct SolveCtx
{
	DBodies B;
	DContacts C;
	DJoint* J;
	int32_t nJ;
	const int32_t* jointIsland;
	const uint32_t* bodyList;
	const uint32_t* bodyStart; // per island, exclusive scan (size islands+1)
	const uint32_t* contactList;
	const uint32_t* contactStart;
	uint32_t* colorMask; // per body
	int32_t* contactColor; // per contact
	m3d_vec3 gravity;
	float h, invH, gravityLen, maxCorrection;
	int32_t iterations;
	int32_t substeps;
};

__device__ __forceinline__ void load_wb(const DBodies& B, int32_t i, WB& b)
{
	float4 P = B.pos[i];
	b.p = f4_v3(P);
	b.invMass = P.w;
	b.q = f4_q(B.rot[i]);
	b.v = f4_v3(B.vel[i]);
	b.w = f4_v3(B.angVel[i]);
	b.iw = load_iw(B, i);
}

__device__ __forceinline__ void store_pose(const DBodies& B, int32_t i, const WB& b)
{
	if (b.invMass == 0.0f)
		return;
	B.pos[i] = v3_f4(b.p, b.invMass);
	B.rot[i] = q_f4(b.q);
}

__device__ __forceinline__ void store_vel(const DBodies& B, int32_t i, const WB& b)
{
	if (b.invMass == 0.0f)
		return;
	float4 V = B.vel[i], W = B.angVel[i];
	B.vel[i] = v3_f4(b.v, V.w);
	B.angVel[i] = v3_f4(b.w, W.w);
}

__device__ void contact_normals_serial(const SolveCtx& X, int32_t c)
{
	int2 ab = X.C.bodies[c];
	float4 PA = X.B.pos[ab.x], PB = X.B.pos[ab.y];
	m3d_vec3 pa = f4_v3(PA), pb = f4_v3(PB);
	m3d_quat qa = f4_q(X.B.rot[ab.x]), qb = f4_q(X.B.rot[ab.y]);
	solve_normals_serial(X.C, c, pa, qa, PA.w, pb, qb, PB.w, X.maxCorrection);
	if (PA.w != 0.0f)
	{
		X.B.pos[ab.x] = v3_f4(pa, PA.w);
		X.B.rot[ab.x] = q_f4(qa);
	}
	if (PB.w != 0.0f)
	{
		X.B.pos[ab.y] = v3_f4(pb, PB.w);
		X.B.rot[ab.y] = q_f4(qb);
	}
}

__device__ void contact_normals_packet(const SolveCtx& X, int32_t c, int32_t maxPoints)
{
	int2 ab = X.C.bodies[c];
	float4 PA = X.B.pos[ab.x], PB = X.B.pos[ab.y];
	m3d_vec3 pa = f4_v3(PA), pb = f4_v3(PB);
	m3d_quat qa = f4_q(X.B.rot[ab.x]), qb = f4_q(X.B.rot[ab.y]);
	solve_normals_packet(X.C, c, maxPoints, pa, qa, PA.w, pb, qb, PB.w, X.maxCorrection);
	if (PA.w > 0.0f)
	{
		X.B.pos[ab.x] = v3_f4(pa, PA.w);
		X.B.rot[ab.x] = q_f4(qa);
	}
	if (PB.w > 0.0f)
	{
		X.B.pos[ab.y] = v3_f4(pb, PB.w);
		X.B.rot[ab.y] = q_f4(qb);
	}
}

__device__ void contact_friction(const SolveCtx& X, int32_t c)
{
	int2 ab = X.C.bodies[c];
	bool boxBox = bf_shape(X.B.flags[ab.x]) == M3D_SHAPE_BOX && bf_shape(X.B.flags[ab.y]) == M3D_SHAPE_BOX;
	if (!boxBox)
		return;
	WB a, b;
	load_wb(X.B, ab.x, a);
	load_wb(X.B, ab.y, b);
	apply_friction(X.C, c, boxBox, a, b);
	store_pose(X.B, ab.x, a);
	store_pose(X.B, ab.y, b);
}

__device__ void contact_velocity(const SolveCtx& X, int32_t c)
{
	int2 ab = X.C.bodies[c];
	WB a, b;
	load_wb(X.B, ab.x, a);
	load_wb(X.B, ab.y, b);
	m3d_vec3 pvA = f4_v3(X.B.prevVel[ab.x]), pwA = f4_v3(X.B.prevAngVel[ab.x]);
	m3d_vec3 pvB = f4_v3(X.B.prevVel[ab.y]), pwB = f4_v3(X.B.prevAngVel[ab.y]);
	VelShape sa, sbv;
	sa.type = bf_shape(X.B.flags[ab.x]);
	sa.radius = X.B.shape[ab.x].w;
	sa.halfHeight = X.B.mat[ab.x].x;
	sbv.type = bf_shape(X.B.flags[ab.y]);
	sbv.radius = X.B.shape[ab.y].w;
	sbv.halfHeight = X.B.mat[ab.y].x;
	solve_velocity(X.C, c, X.h, X.gravityLen, a, b, pvA, pwA, pvB, pwB, sa, sbv);
	store_vel(X.B, ab.x, a);
	store_vel(X.B, ab.y, b);
}

__device__ void joint_op(const SolveCtx& X, int32_t j, int op)
{
	DJoint& jt = X.J[j];
	WB a, b;
	load_wb(X.B, jt.bodyA, a);
	load_wb(X.B, jt.bodyB, b);
	if (op == 0)
	{
		prepare_joint(jt, a, b);
		return;
	}
	if (op == 1)
	{
		solve_joint_position(jt, a, b, X.h);
		store_pose(X.B, jt.bodyA, a);
		store_pose(X.B, jt.bodyB, b);
	}
	else
	{
		solve_joint_velocity(jt, a, b, X.h);
		store_vel(X.B, jt.bodyA, a);
		store_vel(X.B, jt.bodyB, b);
	}
}

// One lane solves one island with the reference's exact serial schedule.
__global__ void k_solve_islands_serial(SolveCtx X, const int32_t* islands, int32_t nIslands)
{
	int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= nIslands)
		return;
	int32_t isl = islands[t];
	uint32_t b0 = X.bodyStart[isl], b1 = X.bodyStart[isl + 1];
	uint32_t c0 = X.contactStart[isl], c1 = X.contactStart[isl + 1];
	bool colored = (int32_t)(c1 - c0) >= kBigIslandContacts;
	bool anyJoint = false;
	for (int32_t j = 0; j < X.nJ; ++j)
		anyJoint |= X.jointIsland[j] == isl;

	if (anyJoint)
	{
		for (uint32_t k = b0; k < b1; ++k)
		{
			int32_t bi = (int32_t)X.bodyList[k];
			m3d_mat3 m = inv_inertia_world(f4_q(X.B.rot[bi]), f4_v3(X.B.invIL[bi]));
			store_iw(X.B, bi, m);
		}
		for (int32_t j = 0; j < X.nJ; ++j)
			if (X.jointIsland[j] == isl)
				joint_op(X, j, 0);
	}

	int32_t colorCount = 0;
	if (colored)
	{
		// greedy first-fit in island contact order; src/xpbd.cpp:1179-1235
		for (uint32_t k = b0; k < b1; ++k)
			X.colorMask[X.bodyList[k]] = 0;
		for (uint32_t k = c0; k < c1; ++k)
		{
			int32_t c = (int32_t)X.contactList[k];
			int2 ab = X.C.bodies[c];
			bool dA = X.B.pos[ab.x].w > 0.0f, dB = X.B.pos[ab.y].w > 0.0f;
			uint32_t used = 0;
			if (dA)
				used |= X.colorMask[ab.x];
			if (dB)
				used |= X.colorMask[ab.y];
			uint32_t fb = ~used;
			int32_t color = fb != 0 ? __builtin_ctz(fb) : kMaxGraphColors;
			if (color >= kMaxGraphColors)
			{
				X.contactColor[c] = kMaxGraphColors;
				continue;
			}
			X.contactColor[c] = color;
			if (dA)
				X.colorMask[ab.x] |= 1u << color;
			if (dB)
				X.colorMask[ab.y] |= 1u << color;
			colorCount = color + 1 > colorCount ? color + 1 : colorCount;
		}
	}

	for (int sub = 0; sub < X.substeps; ++sub)
	{
		for (uint32_t k = b0; k < b1; ++k)
			integrate_body(X.B, (int32_t)X.bodyList[k], X.h, X.gravity);
		for (uint32_t k = c0; k < c1; ++k)
			prepare_contact(X.B, X.C, (int32_t)X.contactList[k], X.h);
		for (int32_t j = 0; j < X.nJ; ++j)
		{
			if (X.jointIsland[j] == isl)
			{
				X.J[j].lambda = 0.0f;
				X.J[j].motorLambda = 0.0f;
			}
		}
		for (int it = 0; it < X.iterations; ++it)
		{
			for (int32_t j = 0; j < X.nJ; ++j)
				if (X.jointIsland[j] == isl)
					joint_op(X, j, 1);
			if (!colored)
			{
				for (uint32_t k = c0; k < c1; ++k)
					contact_normals_serial(X, (int32_t)X.contactList[k]);
				continue;
			}
			for (int32_t color = 0; color < colorCount; ++color)
			{
				// packets of 8 consecutive bucket members
				int32_t lane[8];
				int32_t nl = 0;
				for (uint32_t k = c0; k <= c1; ++k)
				{
					bool end = k == c1;
					if (!end)
					{
						int32_t c = (int32_t)X.contactList[k];
						if (X.contactColor[c] != color)
							continue;
						lane[nl++] = c;
					}
					if (nl == 8 || (end && nl > 0))
					{
						int32_t mp = 0;
						for (int l = 0; l < nl; ++l)
						{
							int32_t pc = f2i(X.C.normal[lane[l]].w);
							mp = pc > mp ? pc : mp;
						}
						for (int l = 0; l < nl; ++l)
							contact_normals_packet(X, lane[l], mp);
						nl = 0;
					}
				}
			}
			for (uint32_t k = c0; k < c1; ++k)
			{
				int32_t c = (int32_t)X.contactList[k];
				if (X.contactColor[c] == kMaxGraphColors)
					contact_normals_serial(X, c);
			}
		}
		if (!colored)
		{
			for (uint32_t k = c0; k < c1; ++k)
				contact_friction(X, (int32_t)X.contactList[k]);
		}
		else
		{
			for (int32_t color = 0; color <= kMaxGraphColors; ++color)
			{
				if (color == colorCount && color < kMaxGraphColors)
					color = kMaxGraphColors;
				for (uint32_t k = c0; k < c1; ++k)
				{
					int32_t c = (int32_t)X.contactList[k];
					if (X.contactColor[c] == color)
						contact_friction(X, c);
				}
			}
		}
		for (uint32_t k = b0; k < b1; ++k)
			reconstruct_velocity(X.B, (int32_t)X.bodyList[k], X.invH);
		if (!colored)
		{
			for (uint32_t k = c0; k < c1; ++k)
				contact_velocity(X, (int32_t)X.contactList[k]);
		}
		else
		{
			for (int32_t color = 0; color <= kMaxGraphColors; ++color)
			{
				if (color == colorCount && color < kMaxGraphColors)
					color = kMaxGraphColors;
				for (uint32_t k = c0; k < c1; ++k)
				{
					int32_t c = (int32_t)X.contactList[k];
					if (X.contactColor[c] == color)
						contact_velocity(X, c);
				}
			}
		}
		for (int32_t j = 0; j < X.nJ; ++j)
			if (X.jointIsland[j] == isl)
				joint_op(X, j, 2);
	}
}

// ------------------------------------------------------------- sleep

__global__ void k_sleep_prep(const uint32_t* flags, const int32_t* islandOf, const uint32_t* islandAwake, int32_t n,
							 int32_t* firstNoSleep)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	int32_t isl = islandOf[i];
	if (isl < 0 || !islandAwake[isl])
		return;
	if (!(flags[i] & BF_ENABLE_SLEEP))
		atomicMin(&firstNoSleep[isl], i);
}

__global__ void k_sleep_update(DBodies B, const int32_t* islandOf, const uint32_t* islandAwake, int32_t n, float dt,
							   const int32_t* firstNoSleep, uint32_t* minSleepBits)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	int32_t isl = islandOf[i];
	if (isl < 0 || !islandAwake[isl] || i >= firstNoSleep[isl])
		return;
	m3d_vec3 v = f4_v3(B.vel[i]), w = f4_v3(B.angVel[i]);
	float4 T = B.iw1[i], E = B.iw2[i];
	float sleepTime = T.w, se = E.w;
	float energy = m3d_length_sq(v) + m3d_length_sq(w);
	se += kSleepEnergyAlpha * (energy - se);
	if (se > kSleepWakeFactor * kSleepEnergyThresh)
		sleepTime = 0.0f;
	else if (se < kSleepEnergyThresh)
		sleepTime += dt;
	else
		sleepTime += 0.5f * dt;
	T.w = sleepTime;
	E.w = se;
	B.iw1[i] = T;
	B.iw2[i] = E;
	atomicMin(&minSleepBits[isl], f2u(sleepTime));
}

__global__ void k_sleep_apply(DBodies B, const int32_t* islandOf, const uint32_t* islandAwake, int32_t n,
							  const int32_t* firstNoSleep, const uint32_t* minSleepBits, uint32_t* counters)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	int32_t isl = islandOf[i];
	if (isl < 0 || !islandAwake[isl])
		return;
	if (firstNoSleep[isl] != 0x7FFFFFFF)
		return;
	if (!(u2f(minSleepBits[isl]) >= kTimeToSleep))
		return;
	B.flags[i] &= ~BF_AWAKE;
	B.sb[i] &= (uint8_t)~SB_AWAKE;
	float4 V = B.vel[i], W = B.angVel[i];
	B.vel[i] = make_float4(0.0f, 0.0f, 0.0f, V.w);
	B.angVel[i] = make_float4(0.0f, 0.0f, 0.0f, W.w);
	counters[CT_STABLE_DIRTY] = 1u;
}

// ---------------------------------------------------------- end of step

__global__ void k_finish(DBodies B, int32_t n, float dt, int clearForces, uint32_t* counters)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	uint8_t f = B.sb[i];
	if ((f & SB_KINEMATIC) && (f & SB_AWAKE))
	{
		float4 P = B.pos[i];
		m3d_vec3 p = m3d_mul_add(f4_v3(P), dt, f4_v3(B.vel[i]));
		m3d_quat q = m3d_quat_integrate(f4_q(B.rot[i]), f4_v3(B.angVel[i]), dt);
		B.pos[i] = v3_f4(p, P.w);
		B.rot[i] = q_f4(q);
	}
	if (clearForces && (f & SB_DYNAMIC))
	{
		B.force[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		B.torque[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	}
	if (f & SB_IN_USE)
	{
		atomicAdd(&counters[CT_BODY_COUNT], 1u);
		if (f & SB_AWAKE)
			atomicAdd(&counters[CT_AWAKE_BODIES], 1u);
	}
}

__global__ void k_mark_flag_list(const uint32_t* flag, int32_t n, const uint32_t* offs, int32_t* out)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n && flag[i])
		out[offs[i]] = i;
}

// ----------------------------------------------- parallel island solve
//
// All awake islands without joints are solved together, one Gauss-Seidel
// sweep per "slot". Within a slot no two constraints share a written
// (invMass > 0) body, so a sweep is a plain parallel launch, and the
// per-body update order equals the reference's:
//   - colored islands (>= 300 contacts): slot = greedy color 0..7, then the
//     serial overflow pass as slots 8 + level (src/xpbd.cpp:1292-1311);
//   - other islands: slot = level of the contact in island order, i.e. one
//     more than the deepest earlier contact sharing a written body
//     (src/xpbd.cpp:947-950).
// The greedy coloring and the levels are computed exactly as least
// fixpoints of their defining recurrences over each body's contact list,
// iterated in place until no value changes.

constexpr uint32_t IM_AWAKE = 1u;
constexpr uint32_t IM_COLORED = 2u;
constexpr uint32_t IM_JOINT = 4u;

__global__ void k_island_modes(const uint32_t* islandAwake, const uint32_t* contactStart, int32_t I, uint32_t* mode,
							   uint32_t* counters)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= I)
		return;
	uint32_t m = islandAwake[i] ? IM_AWAKE : 0u;
	if ((int32_t)(contactStart[i + 1] - contactStart[i]) >= kBigIslandContacts)
	{
		m |= IM_COLORED;
		if (islandAwake[i])
			atomicAdd(&counters[CT_COLORED_ISLANDS], 1u);
	}
	mode[i] = m;
}

__global__ void k_joint_modes(const int32_t* jointIsland, int32_t nJ, uint32_t* mode)
{
	int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j < nJ && jointIsland[j] >= 0)
		atomicOr(&mode[jointIsland[j]], IM_JOINT);
}

// Entries k index the island-sorted contact list; island order equals
// entry order within an island. Adjacency candidates: two per entry.
__global__ void k_adj_emit(const uint32_t* contactList, const uint64_t* entryIsland, const uint32_t* mode,
						   const int2* bodies, const float4* pos, int32_t NC, int32_t n, uint64_t* keys, uint32_t* vals,
						   uint32_t* bodyCount)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= NC)
		return;
	int32_t c = (int32_t)contactList[k];
	uint32_t m = mode[(int32_t)entryIsland[k]];
	int2 ab = bodies[c];
	bool par = !(m & IM_JOINT);
	bool da = par && pos[ab.x].w > 0.0f;
	bool db = par && pos[ab.y].w > 0.0f;
	keys[2 * k] = da ? (uint64_t)ab.x : (uint64_t)n;
	vals[2 * k] = (uint32_t)k;
	keys[2 * k + 1] = db ? (uint64_t)ab.y : (uint64_t)n;
	vals[2 * k + 1] = (uint32_t)k;
	if (da)
		atomicAdd(&bodyCount[ab.x], 1u);
	if (db)
		atomicAdd(&bodyCount[ab.y], 1u);
}

struct OrderCtx
{
	const uint32_t* contactList;
	const uint64_t* entryIsland;
	const uint32_t* mode;
	const int2* bodies;
	const float4* pos;
	const uint32_t* adjStart; // per body, exclusive scan (n + 1)
	const uint32_t* adj;      // entries sorted by (body, entry)
	int32_t NC;
};

__global__ void k_color_round(OrderCtx O, int32_t* color, uint32_t* changed)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= O.NC)
		return;
	uint32_t m = O.mode[(int32_t)O.entryIsland[k]];
	if ((m & (IM_COLORED | IM_JOINT)) != IM_COLORED)
		return;
	int2 ab = O.bodies[O.contactList[k]];
	uint32_t used = 0;
	for (int side = 0; side < 2; ++side)
	{
		int32_t x = side == 0 ? ab.x : ab.y;
		if (!(O.pos[x].w > 0.0f))
			continue;
		for (uint32_t e = O.adjStart[x]; e < O.adjStart[x + 1]; ++e)
		{
			int32_t k2 = (int32_t)O.adj[e];
			if (k2 >= k)
				break;
			int32_t col = __atomic_load_n(&color[k2], __ATOMIC_RELAXED);
			if (col < kMaxGraphColors)
				used |= 1u << col;
		}
	}
	uint32_t fb = ~used;
	int32_t nc = (int32_t)__builtin_ctz(fb);
	nc = nc < kMaxGraphColors ? nc : kMaxGraphColors;
	if (nc != __atomic_load_n(&color[k], __ATOMIC_RELAXED))
	{
		__atomic_store_n(&color[k], nc, __ATOMIC_RELAXED);
		atomicAdd(changed, 1u);
	}
}

__device__ __forceinline__ bool sequenced(const OrderCtx& O, const int32_t* color, int32_t k)
{
	uint32_t m = O.mode[(int32_t)O.entryIsland[k]];
	if (m & IM_JOINT)
		return false;
	return !(m & IM_COLORED) || __atomic_load_n(&color[k], __ATOMIC_RELAXED) == kMaxGraphColors;
}

__global__ void k_level_round(OrderCtx O, const int32_t* color, int32_t* level, uint32_t* changed)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= O.NC || !sequenced(O, color, k))
		return;
	int2 ab = O.bodies[O.contactList[k]];
	int32_t lv = 0;
	for (int side = 0; side < 2; ++side)
	{
		int32_t x = side == 0 ? ab.x : ab.y;
		if (!(O.pos[x].w > 0.0f))
			continue;
		for (uint32_t e = O.adjStart[x]; e < O.adjStart[x + 1]; ++e)
		{
			int32_t k2 = (int32_t)O.adj[e];
			if (k2 >= k)
				break;
			if (sequenced(O, color, k2))
			{
				int32_t l2 = __atomic_load_n(&level[k2], __ATOMIC_RELAXED) + 1;
				lv = l2 > lv ? l2 : lv;
			}
		}
	}
	if (lv != __atomic_load_n(&level[k], __ATOMIC_RELAXED))
	{
		__atomic_store_n(&level[k], lv, __ATOMIC_RELAXED);
		atomicAdd(changed, 1u);
	}
}

// slot key per entry (joint islands excluded with an all-ones key)
__global__ void k_slot_keys(OrderCtx O, const int32_t* color, const int32_t* level, int bitsI, uint64_t* keys,
							uint32_t* vals, int32_t* contactColor, uint32_t* counters)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= O.NC)
		return;
	int32_t isl = (int32_t)O.entryIsland[k];
	uint32_t m = O.mode[isl];
	uint64_t key = ~0ull;
	if (!(m & IM_JOINT))
	{
		int32_t slot;
		if (m & IM_COLORED)
		{
			int32_t col = color[k];
			contactColor[O.contactList[k]] = col;
			slot = col < kMaxGraphColors ? col : kMaxGraphColors + level[k];
		}
		else
		{
			slot = level[k];
		}
		key = ((uint64_t)slot << bitsI) | (uint64_t)isl;
		atomicMax(&counters[CT_SLOTS], (uint32_t)slot + 1u);
		atomicAdd(&counters[CT_PAR_ENTRIES], 1u);
	}
	keys[k] = key;
	vals[k] = O.contactList[k];
}

__global__ void k_slot_starts(const uint64_t* keys, int32_t NP, int bitsI, int32_t nSlots, uint32_t* slotStart)
{
	int32_t s = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (s > nSlots)
		return;
	uint64_t target = (uint64_t)s << bitsI;
	int32_t l = 0, r = NP;
	while (l < r)
	{
		int32_t mid = (l + r) >> 1;
		if (keys[mid] < target)
			l = mid + 1;
		else
			r = mid;
	}
	slotStart[s] = (uint32_t)l;
}

// packet = 8 consecutive members of a (color, island) bucket
__global__ void k_packets(const uint64_t* keys, int32_t NP, int bitsI, const uint32_t* mode, int2* packet)
{
	int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (e >= NP)
		return;
	uint64_t key = keys[e];
	int32_t slot = (int32_t)(key >> bitsI);
	int32_t isl = (int32_t)(key & ((1ull << bitsI) - 1));
	if (!(mode[isl] & IM_COLORED) || slot >= kMaxGraphColors)
	{
		packet[e] = make_int2(-1, -1);
		return;
	}
	int32_t l = 0, r = e;
	while (l < r)
	{
		int32_t mid = (l + r) >> 1;
		if (keys[mid] < key)
			l = mid + 1;
		else
			r = mid;
	}
	int32_t gs = l;
	l = e;
	r = NP;
	while (l < r)
	{
		int32_t mid = (l + r) >> 1;
		if (keys[mid] <= key)
			l = mid + 1;
		else
			r = mid;
	}
	int32_t ge = l;
	int32_t ps = gs + ((e - gs) & ~7);
	int32_t pe = ps + 8 < ge ? ps + 8 : ge;
	packet[e] = make_int2(ps, pe);
}

struct SweepCtx
{
	DBodies B;
	DContacts C;
	const uint32_t* entries; // contact per sorted entry
	const int2* packet;
	float h, gravityLen, maxCorrection;
};

__global__ void __launch_bounds__(256) k_sweep_normals(SweepCtx X, int32_t s0, int32_t s1)
{
	int32_t e = s0 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (e >= s1)
		return;
	int32_t c = (int32_t)X.entries[e];
	int2 pk = X.packet[e];
	int2 ab = X.C.bodies[c];
	float4 PA = X.B.pos[ab.x], PB = X.B.pos[ab.y];
	m3d_vec3 pa = f4_v3(PA), pb = f4_v3(PB);
	m3d_quat qa = f4_q(X.B.rot[ab.x]), qb = f4_q(X.B.rot[ab.y]);
	if (pk.x >= 0)
	{
		int32_t mp = 0;
		for (int32_t q = pk.x; q < pk.y; ++q)
		{
			int32_t pc = f2i(X.C.normal[X.entries[q]].w);
			mp = pc > mp ? pc : mp;
		}
		solve_normals_packet(X.C, c, mp, pa, qa, PA.w, pb, qb, PB.w, X.maxCorrection);
		if (PA.w > 0.0f)
		{
			X.B.pos[ab.x] = v3_f4(pa, PA.w);
			X.B.rot[ab.x] = q_f4(qa);
		}
		if (PB.w > 0.0f)
		{
			X.B.pos[ab.y] = v3_f4(pb, PB.w);
			X.B.rot[ab.y] = q_f4(qb);
		}
		return;
	}
	solve_normals_serial(X.C, c, pa, qa, PA.w, pb, qb, PB.w, X.maxCorrection);
	if (PA.w != 0.0f)
	{
		X.B.pos[ab.x] = v3_f4(pa, PA.w);
		X.B.rot[ab.x] = q_f4(qa);
	}
	if (PB.w != 0.0f)
	{
		X.B.pos[ab.y] = v3_f4(pb, PB.w);
		X.B.rot[ab.y] = q_f4(qb);
	}
}

__global__ void __launch_bounds__(256) k_sweep_friction(SweepCtx X, int32_t s0, int32_t s1)
{
	int32_t e = s0 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (e >= s1)
		return;
	int32_t c = (int32_t)X.entries[e];
	int2 ab = X.C.bodies[c];
	if (bf_shape(X.B.flags[ab.x]) != M3D_SHAPE_BOX || bf_shape(X.B.flags[ab.y]) != M3D_SHAPE_BOX)
		return;
	WB a, b;
	load_wb(X.B, ab.x, a);
	load_wb(X.B, ab.y, b);
	apply_friction(X.C, c, true, a, b);
	store_pose(X.B, ab.x, a);
	store_pose(X.B, ab.y, b);
}

__global__ void __launch_bounds__(256) k_sweep_velocity(SweepCtx X, int32_t s0, int32_t s1)
{
	int32_t e = s0 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (e >= s1)
		return;
	int32_t c = (int32_t)X.entries[e];
	int2 ab = X.C.bodies[c];
	WB a, b;
	load_wb(X.B, ab.x, a);
	load_wb(X.B, ab.y, b);
	m3d_vec3 pvA = f4_v3(X.B.prevVel[ab.x]), pwA = f4_v3(X.B.prevAngVel[ab.x]);
	m3d_vec3 pvB = f4_v3(X.B.prevVel[ab.y]), pwB = f4_v3(X.B.prevAngVel[ab.y]);
	VelShape sa, sbv;
	sa.type = bf_shape(X.B.flags[ab.x]);
	sa.radius = X.B.shape[ab.x].w;
	sa.halfHeight = X.B.mat[ab.x].x;
	sbv.type = bf_shape(X.B.flags[ab.y]);
	sbv.radius = X.B.shape[ab.y].w;
	sbv.halfHeight = X.B.mat[ab.y].x;
	solve_velocity(X.C, c, X.h, X.gravityLen, a, b, pvA, pwA, pvB, pwB, sa, sbv);
	store_vel(X.B, ab.x, a);
	store_vel(X.B, ab.y, b);
}

__global__ void __launch_bounds__(256) k_prepare_entries(DBodies B, DContacts C, const uint32_t* entries, int32_t NP,
														 float h)
{
	int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (e < NP)
		prepare_contact(B, C, (int32_t)entries[e], h);
}

__global__ void __launch_bounds__(256) k_integrate_list(DBodies B, const uint32_t* list, const int32_t* islandOf,
														const uint32_t* mode, int32_t NB, float h, m3d_vec3 gravity)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= NB)
		return;
	int32_t i = (int32_t)list[k];
	if (mode[islandOf[i]] & IM_JOINT)
		return;
	integrate_body(B, i, h, gravity);
}

__global__ void __launch_bounds__(256) k_reconstruct_list(DBodies B, const uint32_t* list, const int32_t* islandOf,
														  const uint32_t* mode, int32_t NB, float invH)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= NB)
		return;
	int32_t i = (int32_t)list[k];
	if (mode[islandOf[i]] & IM_JOINT)
		return;
	reconstruct_velocity(B, i, invH);
}

__global__ void k_joint_island_list(const uint32_t* mode, int32_t I, const uint32_t* offs, int32_t* out,
									uint32_t* flag)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= I)
		return;
	(void)offs;
	(void)out;
	flag[i] = ((mode[i] & (IM_AWAKE | IM_JOINT)) == (IM_AWAKE | IM_JOINT)) ? 1u : 0u;
}

static bool fixpoint(World& w, int which, const OrderCtx& O, int32_t* color, int32_t* level, int32_t* rounds)
{
	hipStream_t s = w.stream;
	const int batch = 4; // rounds per convergence check
	for (int r = 0; r < 1 << 20; r += batch)
	{
		uint32_t* ch = w.counters.p + CT_CHANGED;
		for (int b = 0; b < batch; ++b)
		{
			if (b == batch - 1) // only the last round of a batch decides convergence
				(void)hipMemsetAsync(ch, 0, sizeof(uint32_t), s);
			if (which == 0)
				k_color_round<<<grids(O.NC), kBlockS, 0, s>>>(O, color, ch);
			else
				k_level_round<<<grids(O.NC), kBlockS, 0, s>>>(O, color, level, ch);
		}
		if (!sync_counters(w))
			return false;
		*rounds += batch;
		if (w.hostCounters[CT_CHANGED] == 0)
			return true;
	}
	return true;
}

// Solves every awake island: joint islands one lane each with the serial
// schedule, all others through the slot sweeps.
static bool solve_awake_islands(World& w, float dt, int substeps, int32_t I, int32_t NB, int32_t NC,
								const uint32_t* bodyList, const uint32_t* contactList, const uint32_t* islandBodies,
								const uint32_t* islandContacts, const uint32_t* islandAwake, int bitsI)
{
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	float h = dt / (float)substeps;
	float invH = h > 0.0f ? 1.0f / h : 0.0f;
	float gravityLen = m3d_length(w.def.gravity);
	float maxCorrection = kMaxDepenetrationSpeed * h;
	int32_t iterations = w.def.positionIterations > 0 ? w.def.positionIterations : 2;
	if (!w.islandMode.ensure(I + 1) || !w.sJointList.ensure(I + 1) || !w.s32c.ensure((size_t)n + 2))
		return false;
	uint32_t* mode = w.islandMode.p;
	k_island_modes<<<grids(I), kBlockS, 0, s>>>(islandAwake, islandContacts, I, mode, w.counters.p);
	if (w.djointCount > 0)
	{
		k_joint_modes<<<grids(w.djointCount), kBlockS, 0, s>>>(w.i32e.p, w.djointCount, mode);
		// joint islands: reference serial schedule, one lane per island
		uint32_t* jflag = w.s32c.p;
		uint32_t* joff = w.s32c.p + (I + 1);
		if (!w.s32c.ensure(2 * (size_t)(I + 1)))
			return false;
		jflag = w.s32c.p;
		joff = w.s32c.p + (I + 1);
		k_joint_island_list<<<grids(I), kBlockS, 0, s>>>(mode, I, nullptr, nullptr, jflag);
		scan_exclusive(jflag, joff, I, w.counters.p + CT_JOINT_ISLANDS, w.primScratch.p, s);
		k_mark_flag_list<<<grids(I), kBlockS, 0, s>>>(jflag, I, joff, w.sJointList.p);
		if (!sync_counters(w))
			return false;
		int32_t NJ = (int32_t)w.hostCounters[CT_JOINT_ISLANDS];
		if (NJ > 0)
		{
			if (!w.u32a.ensure((size_t)n + 1))
				return false;
			SolveCtx X;
			X.B = make_body_view(w);
			X.C = make_contact_view(w);
			X.J = w.djoints.p;
			X.nJ = w.djointCount;
			X.jointIsland = w.i32e.p;
			X.bodyList = bodyList;
			X.bodyStart = islandBodies;
			X.contactList = contactList;
			X.contactStart = islandContacts;
			X.colorMask = w.u32a.p;
			X.contactColor = w.contactColor.p;
			X.gravity = w.def.gravity;
			X.h = h;
			X.invH = invH;
			X.gravityLen = gravityLen;
			X.maxCorrection = maxCorrection;
			X.iterations = iterations;
			X.substeps = substeps;
			k_solve_islands_serial<<<(NJ + 63) / 64, 64, 0, s>>>(X, w.sJointList.p, NJ);
		}
	}

	// ---- slot schedule of the parallel path ----
	int32_t NP = 0;
	int32_t nSlots = 0;
	std::vector<uint32_t> slotStart;
	const uint64_t* entryIsland = w.u64c.p; // island of each island-sorted contact entry
	int bitsS = 1;
	if (NC > 0)
	{
		size_t two = 2 * (size_t)NC;
		int bitsN = 1;
		while ((1ll << bitsN) <= (int64_t)n)
			++bitsN;
		if (!w.s64a.ensure(two) || !w.s64b.ensure(two) || !w.s32a.ensure(two) || !w.s32b.ensure(two) ||
			!w.s32c.ensure((size_t)n + 2) || !w.sColor.ensure(NC) || !w.sLevel.ensure(NC) ||
			!w.primScratch.ensure(std::max(scan_scratch_bytes(two + n + 2), sort_scratch_bytes(two)) + 1024))
			return false;
		uint32_t* bodyCount = w.s32c.p;
		launch_fill_u32(bodyCount, 0u, (int64_t)n + 1, s);
		k_adj_emit<<<grids(NC), kBlockS, 0, s>>>(contactList, entryIsland, mode, w.cc.bodies.p, w.bc.pos.p, NC, n,
												 w.s64a.p, w.s32a.p, bodyCount);
		scan_exclusive(bodyCount, bodyCount, (int64_t)n + 1, nullptr, w.primScratch.p, s);
		radix_sort_pairs(w.s64a.p, w.s32a.p, w.s64b.p, w.s32b.p, (int64_t)two, bitsN, w.primScratch.p, s);
		OrderCtx O;
		O.contactList = contactList;
		O.entryIsland = entryIsland;
		O.mode = mode;
		O.bodies = w.cc.bodies.p;
		O.pos = w.bc.pos.p;
		O.adjStart = bodyCount;
		O.adj = w.s32a.p;
		O.NC = NC;
		launch_fill_i32(w.sColor.p, 0, NC, s);
		launch_fill_i32(w.sLevel.p, 0, NC, s);
		int32_t rounds = 0;
		if (!fixpoint(w, 0, O, w.sColor.p, w.sLevel.p, &rounds) || !fixpoint(w, 1, O, w.sColor.p, w.sLevel.p, &rounds))
			return false;
		w.stats.colorRounds = rounds;
		k_slot_keys<<<grids(NC), kBlockS, 0, s>>>(O, w.sColor.p, w.sLevel.p, bitsI, w.s64b.p, w.s32b.p,
												  w.contactColor.p, w.counters.p);
		if (!sync_counters(w))
			return false;
		nSlots = (int32_t)w.hostCounters[CT_SLOTS];
		NP = (int32_t)w.hostCounters[CT_PAR_ENTRIES];
		w.stats.coloredIslands = (int32_t)w.hostCounters[CT_COLORED_ISLANDS];
		while ((1ll << bitsS) <= (int64_t)nSlots)
			++bitsS;
		++bitsS; // keeps the joint-island sentinel key above every slot key
		radix_sort_pairs(w.s64b.p, w.s32b.p, w.s64a.p, w.s32a.p, NC, bitsI + bitsS, w.primScratch.p, s);
		if (!w.sSlotStart.ensure((size_t)nSlots + 1) || !w.sPacket.ensure(std::max(NP, 1)))
			return false;
		if (NP > 0)
		{
			k_slot_starts<<<grids(nSlots + 1), kBlockS, 0, s>>>(w.s64b.p, NP, bitsI, nSlots, w.sSlotStart.p);
			k_packets<<<grids(NP), kBlockS, 0, s>>>(w.s64b.p, NP, bitsI, mode, w.sPacket.p);
			slotStart.resize(nSlots + 1);
			if (!hip_ok(hipMemcpyAsync(slotStart.data(), w.sSlotStart.p, (nSlots + 1) * sizeof(uint32_t),
									   hipMemcpyDeviceToHost, s),
						"slot starts") ||
				!hip_ok(hipStreamSynchronize(s), "slot starts"))
				return false;
		}
	}
	w.stats.sweeps = nSlots;

	// ---- substeps ----
	DBodies B = make_body_view(w);
	DContacts Cv = make_contact_view(w);
	SweepCtx X;
	X.B = B;
	X.C = Cv;
	X.entries = w.s32b.p;
	X.packet = w.sPacket.p;
	X.h = h;
	X.gravityLen = gravityLen;
	X.maxCorrection = maxCorrection;
	for (int sub = 0; sub < substeps; ++sub)
	{
		k_integrate_list<<<grids(NB), kBlockS, 0, s>>>(B, bodyList, w.islandOf.p, mode, NB, h, w.def.gravity);
		if (NP > 0)
			k_prepare_entries<<<grids(NP), kBlockS, 0, s>>>(B, Cv, w.s32b.p, NP, h);
		for (int it = 0; it < iterations; ++it)
		{
			for (int32_t sl = 0; sl < nSlots; ++sl)
			{
				int32_t a = (int32_t)slotStart[sl], b = (int32_t)slotStart[sl + 1];
				if (b > a)
					k_sweep_normals<<<grids(b - a), kBlockS, 0, s>>>(X, a, b);
			}
		}
		for (int32_t sl = 0; sl < nSlots; ++sl)
		{
			int32_t a = (int32_t)slotStart[sl], b = (int32_t)slotStart[sl + 1];
			if (b > a)
				k_sweep_friction<<<grids(b - a), kBlockS, 0, s>>>(X, a, b);
		}
		k_reconstruct_list<<<grids(NB), kBlockS, 0, s>>>(B, bodyList, w.islandOf.p, mode, NB, invH);
		for (int32_t sl = 0; sl < nSlots; ++sl)
		{
			int32_t a = (int32_t)slotStart[sl], b = (int32_t)slotStart[sl + 1];
			if (b > a)
				k_sweep_velocity<<<grids(b - a), kBlockS, 0, s>>>(X, a, b);
		}
	}
	return true;
}

bool islands_and_solve(World& w, float dt, int substeps)
{
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	int32_t C = w.contactCount;
	w.lastHadIslands = false;
	w.lastIslandCount = 0;
	if (n == 0)
		return true;
	k_any_awake_dyn<<<grids(n), kBlockS, 0, s>>>(w.bc.sb.p, n, w.counters.p);
	if (!sync_counters(w))
		return false;
	if (!w.hostCounters[CT_ANY_AWAKE_DYN])
		return true;

	int64_t m = std::max<int64_t>(n, C);
	if (!w.i32a.ensure(n) || !w.u32a.ensure(m + 1) || !w.u32b.ensure(m + 1) || !w.islandOf.ensure(n) ||
		!w.primScratch.ensure(std::max(scan_scratch_bytes(m + 1), sort_scratch_bytes(m + 1)) + 1024))
		return false;
	int32_t* parent = w.i32a.p;
	k_uf_init<<<grids(n), kBlockS, 0, s>>>(parent, n);
	if (C > 0)
		k_uf_contacts<<<grids(C), kBlockS, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, w.bc.sb.p, C, parent);
	if (w.djointCount > 0)
		k_uf_joints<<<grids(w.djointCount), kBlockS, 0, s>>>(w.djoints.p, w.djointCount, w.bc.flags.p, parent);
	k_uf_roots<<<grids(n), kBlockS, 0, s>>>(parent, w.bc.sb.p, n, w.u32a.p);
	scan_exclusive(w.u32a.p, w.u32b.p, n, w.counters.p + CT_ISLANDS, w.primScratch.p, s);
	if (!sync_counters(w))
		return false;
	int32_t I = (int32_t)w.hostCounters[CT_ISLANDS];
	w.lastHadIslands = true;
	w.lastIslandCount = I;
	if (I == 0)
		return true;
	if (!w.islandAwakeD.ensure(I + 1) || !w.u32d.ensure(2 * (size_t)(I + 1)))
		return false;
	uint32_t* islandAwake = w.islandAwakeD.p;
	launch_fill_u32(islandAwake, 0u, I, s);
	k_island_of<<<grids(n), kBlockS, 0, s>>>(parent, w.bc.sb.p, w.u32b.p, n, w.islandOf.p, islandAwake);
	k_island_wake<<<grids(n), kBlockS, 0, s>>>(make_body_view(w), n, w.islandOf.p, islandAwake, w.counters.p);

	// island body lists (ascending) and contact lists (ascending)
	uint32_t* islandBodies = w.u32d.p;          // counts, then exclusive starts
	uint32_t* islandContacts = w.u32d.p + I + 1;
	launch_fill_u32(islandBodies, 0u, 2 * (int64_t)(I + 1), s);
	int bitsI = 1;
	while ((1ll << bitsI) <= I)
		++bitsI;
	if (!w.u64a.ensure(m) || !w.u64b.ensure(m) || !w.u64c.ensure(m) || !w.sortValsAlt.ensure(m) ||
		!w.i32b.ensure(n) || !w.i32c.ensure(C + 1) || !w.i32d.ensure(m) ||
		!w.i32e.ensure(std::max<int32_t>(w.djointCount, 1)) || !w.contactColor.ensure(std::max<int32_t>(C, 1)))
		return false;
	uint32_t* flag = w.u32a.p;
	uint32_t* offs = w.u32b.p;
	uint32_t* bodyListVals = (uint32_t*)w.i32b.p;
	uint32_t* contactListVals = (uint32_t*)w.i32c.p;
	// bodies
	k_body_island_keys<<<grids(n), kBlockS, 0, s>>>(w.islandOf.p, islandAwake, n, w.u64a.p, (uint32_t*)w.i32d.p,
													 flag, islandBodies);
	scan_exclusive(flag, offs, n, w.counters.p + CT_INTEGRATED, w.primScratch.p, s);
	k_compact_pairs<<<grids(n), kBlockS, 0, s>>>(w.u64a.p, (uint32_t*)w.i32d.p, flag, offs, n, w.u64b.p,
												  bodyListVals);
	// contacts
	launch_fill_i32(w.contactColor.p, -1, C, s);
	if (C > 0)
	{
		k_contact_island<<<grids(C), kBlockS, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, w.cc.normal.p, w.bc.sb.p,
													   w.islandOf.p, islandAwake, C, w.u64a.p, (uint32_t*)w.i32d.p,
													   flag, islandContacts, w.counters.p);
		scan_exclusive(flag, offs, C, w.counters.p + CT_SEG_COUNT, w.primScratch.p, s);
		k_compact_pairs<<<grids(C), kBlockS, 0, s>>>(w.u64a.p, (uint32_t*)w.i32d.p, flag, offs, C, w.u64c.p,
													  contactListVals);
	}
	if (!sync_counters(w))
		return false;
	int32_t NB = (int32_t)w.hostCounters[CT_INTEGRATED];
	int32_t NC = C > 0 ? (int32_t)w.hostCounters[CT_SEG_COUNT] : 0;
	radix_sort_pairs(w.u64b.p, bodyListVals, w.u64a.p, w.sortValsAlt.p, NB, bitsI, w.primScratch.p, s);
	if (NC > 0)
		radix_sort_pairs(w.u64c.p, contactListVals, w.u64a.p, w.sortValsAlt.p, NC, bitsI, w.primScratch.p, s);
	scan_exclusive(islandBodies, islandBodies, I + 1, nullptr, w.primScratch.p, s);
	scan_exclusive(islandContacts, islandContacts, I + 1, nullptr, w.primScratch.p, s);

	// awake island list
	if (!w.u32a.ensure(I + 1))
		return false;
	scan_exclusive(islandAwake, w.u32a.p, I, w.counters.p + CT_AWAKE_ISLANDS, w.primScratch.p, s);
	int32_t* awakeList = w.i32a.p; // parent no longer needed
	k_mark_flag_list<<<grids(I), kBlockS, 0, s>>>(islandAwake, I, w.u32a.p, awakeList);
	if (w.djointCount > 0)
		k_joint_island<<<grids(w.djointCount), kBlockS, 0, s>>>(w.djoints.p, w.djointCount, w.bc.flags.p,
																 w.islandOf.p, islandAwake, w.i32e.p);
	if (!sync_counters(w))
		return false;
	int32_t NA = (int32_t)w.hostCounters[CT_AWAKE_ISLANDS];
	w.stats.awakeDynamicBodies = NB;
	w.stats.solvedContacts = NC;
	w.stats.solvedPoints = (int32_t)w.hostCounters[CT_SOLVED_POINTS];
	w.stats.boxBoxPoints = (int32_t)w.hostCounters[CT_BB_POINTS];

	if (NA > 0 && !solve_awake_islands(w, dt, substeps, I, NB, NC, bodyListVals, contactListVals, islandBodies,
										islandContacts, islandAwake, bitsI))
		return false;
	// sleeping
	if (!w.i32b.ensure(I) || !w.u32b.ensure(I))
		return false;
	launch_fill_i32(w.i32b.p, 0x7FFFFFFF, I, s);
	launch_fill_u32(w.u32b.p, f2u(FLT_MAX), I, s);
	k_sleep_prep<<<grids(n), kBlockS, 0, s>>>(w.bc.flags.p, w.islandOf.p, islandAwake, n, w.i32b.p);
	k_sleep_update<<<grids(n), kBlockS, 0, s>>>(make_body_view(w), w.islandOf.p, islandAwake, n, dt, w.i32b.p,
												w.u32b.p);
	k_sleep_apply<<<grids(n), kBlockS, 0, s>>>(make_body_view(w), w.islandOf.p, islandAwake, n, w.i32b.p, w.u32b.p,
											   w.counters.p);
	return true;
}

void finish_step(World& w, float dt)
{
	int32_t n = w.deviceBodyCount;
	if (n > 0)
		k_finish<<<grids(n), kBlockS, 0, w.stream>>>(make_body_view(w), n, dt, w.anyForcesApplied ? 1 : 0,
													 w.counters.p);
}

} // namespace mx
