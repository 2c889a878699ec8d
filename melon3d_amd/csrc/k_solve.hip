// Islands, sleeping and the end-of-step body passes; the island solve
// itself is in k_isolve.hip.
//
//   islands   lock-free union-find (hooking always attaches the larger root
//             under the smaller, so every root is its component's smallest
//             body index) over touching non-sensor dynamic-dynamic contacts
//             and dynamic-dynamic joints; islands numbered in order of their
//             smallest body, as the reference's ascending scan assigns them
//             (src/world.cpp:1868-1958)
//   lists     island body/contact lists by a stable sort on the island id,
//             so each list stays ascending (src/world.cpp:1974-2010)
//   modes     per island: awake, colored (>= 300 contacts), giant (too many
//             contacts for one workgroup's LDS schedule), has joints
//   sleep     low-pass energy, timers and whole-island sleep
//             (src/world.cpp:2060-2108)
#include "mx_steps.h"
#include "mx_wave.h"

#include <algorithm>

namespace mx
{

DBodies make_body_view(World& w);
DContacts make_contact_view(World& w);

constexpr int kBlockS = 256;
static inline unsigned grids(int64_t n) { return (unsigned)((n + kBlockS - 1) / kBlockS); }

// ----------------------------------------------------------- islands

// Flags an awake dynamic body in the counters; with parent set, also starts
// the union-find forest (every body its own root).
__global__ void k_any_awake_dyn(const uint8_t* sb, int32_t n, uint32_t* counters, int32_t* parent)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	if ((sb[i] & (SB_DYNAMIC | SB_AWAKE)) == (SB_DYNAMIC | SB_AWAKE))
		counters[CT_ANY_AWAKE_DYN] = 1u;
	if (parent != nullptr)
		parent[i] = i;
}

__device__ __forceinline__ int32_t uf_find(int32_t* parent, int32_t x)
{
	for (;;)
	{
		// agent scope, as the CAS in uf_union: the forest lives in device memory
		// and only this kernel's lanes read it
		int32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (p == x)
			return x;
		int32_t gp = __hip_atomic_load(&parent[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (gp != p)
			__hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); // halving: still an ancestor
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

// Batched worlds: one workgroup per sub-world does its union-find in LDS (a
// sub-world's bodies and contacts are contiguous ranges, and no contact
// joins two sub-worlds). Same result as the forest start in k_any_awake_dyn, k_uf_contacts and
// k_uf_roots: each body's root is its island's smallest body.
constexpr int32_t kUfLocalBodies = 4096;

__device__ __forceinline__ int32_t uf_find_lds(int32_t* par, int32_t x)
{
	for (;;)
	{
		int32_t p = __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		if (p == x)
			return x;
		int32_t gp = __hip_atomic_load(&par[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		if (gp != p)
			__hip_atomic_store(&par[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		x = p;
	}
}

__global__ void __launch_bounds__(256) k_uf_sub(const int2* bodies, const float4* mat, const uint8_t* sb,
												const int32_t* subBodyStart, const int32_t* subContactStart,
												int32_t* parent, uint32_t* rootFlag)
{
	__shared__ int32_t par[kUfLocalBodies];
	const int32_t sw = (int32_t)blockIdx.x;
	const int32_t b0 = subBodyStart[sw], nb = subBodyStart[sw + 1] - b0;
	const int32_t c0 = subContactStart[sw], c1 = subContactStart[sw + 1];
	if (nb > kUfLocalBodies)
		return; // the host sends such batches to the global kernels
	for (int32_t i = (int32_t)threadIdx.x; i < nb; i += (int32_t)blockDim.x)
		par[i] = i;
	__syncthreads();
	for (int32_t c = c0 + (int32_t)threadIdx.x; c < c1; c += (int32_t)blockDim.x)
	{
		uint32_t cf = f2u(mat[c].w);
		if (!(cf & CF_TOUCHING) || (cf & CF_SENSOR))
			continue;
		int2 ab = bodies[c];
		int32_t a = ab.x - b0, b = ab.y - b0;
		if ((uint32_t)a >= (uint32_t)nb || (uint32_t)b >= (uint32_t)nb)
			continue;
		if (!((sb[ab.x] & SB_DYNAMIC) && (sb[ab.y] & SB_DYNAMIC)))
			continue;
		for (;;)
		{
			int32_t ra = uf_find_lds(par, a);
			int32_t rb = uf_find_lds(par, b);
			if (ra == rb)
				break;
			int32_t lo = ra < rb ? ra : rb;
			int32_t hi = ra < rb ? rb : ra;
			if (atomicCAS(&par[hi], hi, lo) == hi)
				break;
		}
	}
	__syncthreads();
	for (int32_t i = (int32_t)threadIdx.x; i < nb; i += (int32_t)blockDim.x)
	{
		int32_t r = uf_find_lds(par, i);
		parent[b0 + i] = b0 + r;
		rootFlag[b0 + i] = ((sb[b0 + i] & SB_DYNAMIC) && r == i) ? 1u : 0u;
	}
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
	mark_sub_stable_dirty(B, i);
}

// solver contacts: touching, non-sensor, dynamic body's island awake. Runs
// on reduce_grid workgroups; every lane takes the same number of turns (the
// wave-wide helpers need the whole wave).
__global__ void k_contact_island(const int2* bodies, const float4* mat, const float4* normal, const uint8_t* sb,
								 const int32_t* islandOf, const uint32_t* islandAwake, int32_t C, uint64_t* keys,
								 uint32_t* vals, uint32_t* flag, uint32_t* islandContacts, uint32_t* counters)
{
	uint32_t nTouch = 0, nSolved = 0, nPoints = 0, nFlatPoints = 0;
	for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < C; base += (int64_t)gridDim.x * blockDim.x)
	{
		int32_t c = (int32_t)(base + threadIdx.x);
		bool valid = c < C;
		uint32_t cf = valid ? f2u(mat[c].w) : 0u;
		uint32_t f = 0, pc = 0;
		int32_t isl = -1;
		if (cf & CF_TOUCHING)
		{
			if (!(cf & CF_SENSOR))
			{
				int2 ab = bodies[c];
				int32_t dyn = (sb[ab.x] & SB_DYNAMIC) ? ab.x : ab.y;
				isl = islandOf[dyn];
				if (isl >= 0 && islandAwake[isl])
				{
					f = 1;
					pc = (uint32_t)f2i(normal[c].w);
				}
			}
		}
		nTouch += (cf & CF_TOUCHING) ? 1u : 0u;
		nSolved += f;
		nPoints += pc;
		nFlatPoints += (cf & CF_CURVED) ? 0u : pc;
		wave_add_keyed(islandContacts, f ? (uint32_t)isl : ~0u, 1u);
		if (valid)
		{
			flag[c] = f;
			keys[c] = f ? (uint64_t)isl : 0ull;
			vals[c] = (uint32_t)c;
		}
	}
	block_add<4>(counters, { CT_TOUCHING, CT_SOLVED_CONTACTS, CT_SOLVED_POINTS, CT_BB_POINTS },
				 { nTouch, nSolved, nPoints, nFlatPoints });
}

__global__ void k_body_island_keys(const int32_t* islandOf, const uint32_t* islandAwake, int32_t n, uint64_t* keys,
								   uint32_t* vals, uint32_t* flag, uint32_t* islandBodies)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	bool valid = i < n;
	int32_t isl = valid ? islandOf[i] : -1;
	uint32_t f = (isl >= 0 && islandAwake[isl]) ? 1u : 0u;
	wave_add_keyed(islandBodies, f ? (uint32_t)isl : ~0u, 1u);
	if (!valid)
		return;
	flag[i] = f;
	keys[i] = f ? (uint64_t)isl : 0ull;
	vals[i] = (uint32_t)i;
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


// ------------------------------------------------------------- sleep

__global__ void k_sleep_prep(const uint32_t* flags, const int32_t* islandOf, const uint32_t* islandAwake, int32_t n,
							 int32_t* firstNoSleep)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	int32_t isl = i < n ? islandOf[i] : -1;
	bool hit = isl >= 0 && islandAwake[isl] && !(flags[i] & BF_ENABLE_SLEEP);
	wave_min_keyed((uint32_t*)firstNoSleep, hit ? (uint32_t)isl : ~0u, (uint32_t)i);
}

__global__ void k_sleep_update(DBodies B, const int32_t* islandOf, const uint32_t* islandAwake, int32_t n, float dt,
							   const int32_t* firstNoSleep, uint32_t* minSleepBits)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	int32_t isl = i < n ? islandOf[i] : -1;
	bool live = isl >= 0 && islandAwake[isl] && i < firstNoSleep[isl];
	uint32_t key = ~0u, bits = 0u;
	if (live)
	{
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
		key = (uint32_t)isl;
		bits = f2u(sleepTime); // non-negative: orders like the float
	}
	wave_min_keyed(minSleepBits, key, bits);
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
	mark_sub_stable_dirty(B, i);
}

// ---------------------------------------------------------- end of step

// Runs on reduce_grid workgroups.
__global__ void k_finish(DBodies B, int32_t n, float dt, int clearForces, uint32_t* counters)
{
	uint32_t nUse = 0, nAwake = 0;
	for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int32_t)(gridDim.x * blockDim.x))
	{
		uint8_t f = B.sb[i];
		nUse += (f & SB_IN_USE) ? 1u : 0u;
		nAwake += ((f & SB_IN_USE) && (f & SB_AWAKE)) ? 1u : 0u;
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
	}
	block_add<2>(counters, { CT_BODY_COUNT, CT_AWAKE_BODIES }, { nUse, nAwake });
}

// ---------------------------------------------------------- island modes

// Mode bits from the island's awake flag and its solver-contact count.
__global__ void k_island_modes(const uint32_t* islandAwake, const uint32_t* islandContacts, int32_t I, uint32_t* mode,
							   uint32_t* counters)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= I)
		return;
	uint32_t m = 0;
	if (islandAwake[i])
	{
		// joint islands get IM_JOINT afterwards and leave the grid path
		m = IM_AWAKE | IM_GRID;
		if ((int32_t)islandContacts[i] >= kBigIslandContacts)
		{
			m |= IM_COLORED;
			wave_inc(&counters[CT_COLORED_ISLANDS]);
		}
	}
	mode[i] = m;
}

__global__ void k_joint_modes(const int32_t* jointIsland, int32_t nJ, uint32_t* mode)
{
	int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j < nJ && jointIsland[j] >= 0)
		atomicOr(&mode[jointIsland[j]], IM_JOINT);
}

// ------------------------------------------------------- orchestration

// Islands, solve and sleep with two host synchronizations: one for the
// island count, one for the list sizes.
bool islands_and_solve(World& w, float dt, int substeps)
{
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	int32_t C = w.contactCount;
	w.lastHadIslands = false;
	w.lastIslandCount = 0;
	if (n == 0)
		return true;

	int64_t m = std::max<int64_t>(n, C);
	if (!w.i32a.ensure(n) || !w.u32a.ensure(m + 1) || !w.u32b.ensure(m + 1) || !w.islandOf.ensure(n) ||
		!ensure_prim_scratch(w, m + 1))
		return false;
	int32_t* parent = w.i32a.p;
	const bool ufLocal = w.subCount > 0 && w.djointCount == 0 && w.maxSubBodies <= kUfLocalBodies &&
						 w.subBodyStart.size() == (size_t)w.subCount + 1 && w.subBodyStart[w.subCount] == n;
	k_any_awake_dyn<<<grids(n), kBlockS, 0, s>>>(w.bc.sb.p, n, w.counters.p, ufLocal ? nullptr : parent);
	if (ufLocal)
	{
		k_uf_sub<<<(unsigned)w.subCount, 256, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, w.bc.sb.p, w.subBodyStartD.p,
													  w.subContactStart.p, parent, w.u32a.p);
	}
	else
	{
		if (C > 0)
			k_uf_contacts<<<grids(C), kBlockS, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, w.bc.sb.p, C, parent);
		if (w.djointCount > 0)
			k_uf_joints<<<grids(w.djointCount), kBlockS, 0, s>>>(w.djoints.p, w.djointCount, w.bc.flags.p, parent);
		k_uf_roots<<<grids(n), kBlockS, 0, s>>>(parent, w.bc.sb.p, n, w.u32a.p);
	}
	scan_exclusive(w.u32a.p, w.u32b.p, n, w.counters.p + CT_ISLANDS, w.scanStatus.p, s);
	if (!sync_counters(w) || !queue_transitions(w))
		return false;
	// no awake dynamic body: the reference skips island building entirely
	if (!w.hostCounters[CT_ANY_AWAKE_DYN])
		return true;
	int32_t I = (int32_t)w.hostCounters[CT_ISLANDS];
	w.lastHadIslands = true;
	w.lastIslandCount = I;
	if (I == 0)
		return true;
	if (!w.islandAwakeD.ensure(I + 1) || !w.u32d.ensure(2 * (size_t)(I + 1)) || !w.islandMode.ensure(I + 1))
		return false;
	uint32_t* islandAwake = w.islandAwakeD.p;
	// island body lists (ascending) and contact lists (ascending)
	uint32_t* islandBodies = w.u32d.p; // counts, then exclusive starts
	uint32_t* islandContacts = w.u32d.p + I + 1;
	{
		FillSegs F;
		F.add(islandAwake, 0u, I);
		F.add(islandBodies, 0u, 2 * (int64_t)(I + 1));
		launch_fill_multi(F, s);
	}
	k_island_of<<<grids(n), kBlockS, 0, s>>>(parent, w.bc.sb.p, w.u32b.p, n, w.islandOf.p, islandAwake);
	k_island_wake<<<grids(n), kBlockS, 0, s>>>(make_body_view(w), n, w.islandOf.p, islandAwake, w.counters.p);

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
	k_body_island_keys<<<grids(n), kBlockS, 0, s>>>(w.islandOf.p, islandAwake, n, w.u64a.p, (uint32_t*)w.i32d.p,
													 flag, islandBodies);
	scan_exclusive(flag, offs, n, w.counters.p + CT_INTEGRATED, w.scanStatus.p, s);
	k_compact_pairs<<<grids(n), kBlockS, 0, s>>>(w.u64a.p, (uint32_t*)w.i32d.p, flag, offs, n, w.u64b.p,
												  bodyListVals);
	launch_fill_i32(w.contactColor.p, -1, C, s);
	if (C > 0)
	{
		k_contact_island<<<reduce_grid(C, kBlockS), kBlockS, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, w.cc.normal.p, w.bc.sb.p,
													   w.islandOf.p, islandAwake, C, w.u64a.p, (uint32_t*)w.i32d.p,
													   flag, islandContacts, w.counters.p);
		scan_exclusive(flag, offs, C, w.counters.p + CT_SEG_COUNT, w.scanStatus.p, s);
		k_compact_pairs<<<grids(C), kBlockS, 0, s>>>(w.u64a.p, (uint32_t*)w.i32d.p, flag, offs, C, w.u64c.p,
													  contactListVals);
	}
	uint32_t* mode = w.islandMode.p;
	k_island_modes<<<grids(I), kBlockS, 0, s>>>(islandAwake, islandContacts, I, mode, w.counters.p);
	if (w.djointCount > 0)
	{
		k_joint_island<<<grids(w.djointCount), kBlockS, 0, s>>>(w.djoints.p, w.djointCount, w.bc.flags.p,
																 w.islandOf.p, islandAwake, w.i32e.p);
		k_joint_modes<<<grids(w.djointCount), kBlockS, 0, s>>>(w.i32e.p, w.djointCount, mode);
	}
	if (!sync_counters(w))
		return false;
	int32_t NB = (int32_t)w.hostCounters[CT_INTEGRATED];
	int32_t NC = C > 0 ? (int32_t)w.hostCounters[CT_SEG_COUNT] : 0;
	w.stats.awakeDynamicBodies = NB;
	w.stats.solvedContacts = NC;
	w.stats.solvedPoints = (int32_t)w.hostCounters[CT_SOLVED_POINTS];
	w.stats.boxBoxPoints = (int32_t)w.hostCounters[CT_BB_POINTS];
	w.stats.coloredIslands = (int32_t)w.hostCounters[CT_COLORED_ISLANDS];
	// Odd pass counts leave the lists in the ping-pong buffers (no copy
	// back): the body sort's in sortValsAlt, the contact sort's in u64a and
	// i32d (both free by then: the body keys are not read again, and i32d
	// only fed the compaction).
	uint64_t* entryIsland = w.u64c.p;
	if (radix_sort_pairs_in_place_or_alt(w.u64b.p, bodyListVals, w.u64a.p, w.sortValsAlt.p, NB, bitsI, w.primScratch.p,
										 w.scanStatus.p, s))
		bodyListVals = w.sortValsAlt.p;
	if (NC > 0 && radix_sort_pairs_in_place_or_alt(w.u64c.p, contactListVals, w.u64a.p, (uint32_t*)w.i32d.p, NC, bitsI,
												   w.primScratch.p, w.scanStatus.p, s))
	{
		entryIsland = w.u64a.p;
		contactListVals = (uint32_t*)w.i32d.p;
	}
	scan_exclusive2(islandBodies, islandBodies, nullptr, islandContacts, islandContacts, nullptr, I + 1, w.scanStatus.p,
					s);

	if (NB > 0)
	{
		IslandLists L;
		L.bodyList = bodyListVals;
		L.bodyStart = islandBodies;
		L.contactList = contactListVals;
		L.contactStart = islandContacts;
		L.entryIsland = entryIsland;
		L.mode = mode;
		L.islandOf = w.islandOf.p;
		L.jointIsland = w.i32e.p;
		L.I = I;
		L.NB = NB;
		L.NC = NC;
		L.bitsI = bitsI;
		if (!diff_record_step(w, L, dt, substeps) || !solve_islands(w, L, dt, substeps))
			return false;
	}

	// sleeping
	if (!w.i32d.ensure(I) || !w.u32a.ensure(I))
		return false;
	{
		FillSegs F;
		F.add((uint32_t*)w.i32d.p, 0x7FFFFFFFu, I);
		F.add(w.u32a.p, f2u(FLT_MAX), I);
		launch_fill_multi(F, s);
	}
	k_sleep_prep<<<grids(n), kBlockS, 0, s>>>(w.bc.flags.p, w.islandOf.p, islandAwake, n, w.i32d.p);
	k_sleep_update<<<grids(n), kBlockS, 0, s>>>(make_body_view(w), w.islandOf.p, islandAwake, n, dt, w.i32d.p,
												w.u32a.p);
	k_sleep_apply<<<grids(n), kBlockS, 0, s>>>(make_body_view(w), w.islandOf.p, islandAwake, n, w.i32d.p, w.u32a.p,
											   w.counters.p);
	return true;
}

void finish_step(World& w, float dt)
{
	int32_t n = w.deviceBodyCount;
	if (n > 0)
		k_finish<<<reduce_grid(n, kBlockS), kBlockS, 0, w.stream>>>(make_body_view(w), n, dt, w.anyForcesApplied ? 1 : 0,
													 w.counters.p);
}

} // namespace mx
