// Island solve kernels. Every awake island is solved on one cooperative
// grid. Each contact gets a slot: colored islands (>= 300 contacts) their
// greedy first-fit color 0..7 (src/xpbd.cpp:1179-1235), then 8 + the
// serial-order level of their overflow pass; other islands 8 + the
// serial-order level of their contacts. Colors and levels are least
// fixpoints of their recurrences, iterated in place until a round changes
// nothing; then the substep loop sweeps the slots with a grid barrier
// between slots.
//
// A slot never writes one body twice, so each slot is a plain parallel
// pass, and every body sees its constraints in the reference's order:
// colors 0..7 as the colored path's buckets (packet arithmetic, 8
// consecutive bucket members per packet), the overflow and the small
// islands in serial contact order (serial arithmetic). Islands are
// body-disjoint, so sweeping all of them slot by slot together changes
// nothing for any one of them.
//
// Joints (src/xpbd.cpp:413-581) run where the reference runs them in every
// island (:913-921, :938-941, :966-969 serial; :1288-1291, :1364-1367
// colored): prepared once per step after the island's inverse inertias are
// refreshed, their impulses cleared each substep, solved before the
// contacts in every position iteration and after them in the velocity pass.
// Their serial order is split into levels (a joint's level is one more than
// the highest level of the earlier joints on one of its dynamic bodies), one
// launch per level, so joints of one level touch disjoint bodies.
#include "mx_steps.h"
#include "mx_wave.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>

namespace mx
{

DBodies make_body_view(World& w);
DContacts make_contact_view(World& w);

static inline unsigned gridk(int64_t n) { return (unsigned)((n + 255) / 256); }

// --------------------------------------------------------- contact passes

struct PassCtx
{
	DBodies B;
	DContacts C;
	float h, gravityLen, maxCorrection;
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

// Solve record of one scheduled contact: contact, body A, body B, and its
// form: bits 0-7 the packet's maxPoints (kRecSerial: serial form), bit 8 set
// for box-box pairs (the only pairs with stiction friction); the grid-wide
// solve's records also hold the contact's point count in bits 12-14 and set
// bit 15 for curved pairs. Building it once per substep leaves one dependent
// load in front of each pass.
constexpr int32_t kRecSerial = 0xFF;
constexpr int32_t kRecBoxBox = 0x100;
constexpr int kRecCountShift = 12;
constexpr int32_t kRecCurved = 0x8000;

__device__ __forceinline__ int rec_points(int4 r) { return (r.w >> kRecCountShift) & 7; }

__device__ __forceinline__ int4 make_rec(const PassCtx& X, int32_t c, int32_t maxPoints)
{
	int2 ab = X.C.bodies[c];
	bool bb = bf_shape(X.B.flags[ab.x]) == M3D_SHAPE_BOX && bf_shape(X.B.flags[ab.y]) == M3D_SHAPE_BOX;
	return make_int4(c, ab.x, ab.y, (maxPoints < 0 ? kRecSerial : maxPoints) | (bb ? kRecBoxBox : 0));
}

// The three Gauss-Seidel passes over scheduled contacts. load() fetches the
// contact-side inputs, which only the contact itself writes, so a sweep may
// issue it a slot early; run() reads the bodies and solves.
struct NormalsPass
{
	using In = NormalsIn;
	__device__ static void load(const PassCtx& X, int4 r, In& d) { load_normals_in(X.C, r.x, d); }
	__device__ static void run(const PassCtx& X, int4 r, const In& d)
	{
		float4 PA = X.B.pos[r.y], PB = X.B.pos[r.z];
		m3d_vec3 pa = f4_v3(PA), pb = f4_v3(PB);
		m3d_quat qa = f4_q(X.B.rot[r.y]), qb = f4_q(X.B.rot[r.z]);
		int32_t mp = r.w & 0xFF;
		bool wa, wb;
		if (mp == kRecSerial)
		{
			solve_normals_serial_in(X.C, r.x, d, pa, qa, PA.w, pb, qb, PB.w, X.maxCorrection);
			wa = PA.w != 0.0f;
			wb = PB.w != 0.0f;
		}
		else
		{
			solve_normals_packet_in(X.C, r.x, d, mp, pa, qa, PA.w, pb, qb, PB.w, X.maxCorrection);
			wa = PA.w > 0.0f;
			wb = PB.w > 0.0f;
		}
		if (wa)
		{
			X.B.pos[r.y] = v3_f4(pa, PA.w);
			X.B.rot[r.y] = q_f4(qa);
		}
		if (wb)
		{
			X.B.pos[r.z] = v3_f4(pb, PB.w);
			X.B.rot[r.z] = q_f4(qb);
		}
	}
};

struct FrictionPass
{
	using In = FrictionIn;
	__device__ static void load(const PassCtx& X, int4 r, In& d)
	{
		if (r.w & kRecBoxBox)
			load_friction_in(X.C, r.x, d);
	}
	__device__ static void run(const PassCtx& X, int4 r, const In& d)
	{
		if (!(r.w & kRecBoxBox))
			return;
		WB a, b;
		load_wb(X.B, r.y, a);
		load_wb(X.B, r.z, b);
		apply_friction_in(X.C, r.x, d, a, b);
		store_pose(X.B, r.y, a);
		store_pose(X.B, r.z, b);
	}
};

struct VelocityPass
{
	using In = VelocityIn;
	__device__ static void load(const PassCtx& X, int4 r, In& d) { load_velocity_in(X.C, r.x, d); }
	__device__ static void run(const PassCtx& X, int4 r, const In& d)
	{
		WB a, b;
		load_wb(X.B, r.y, a);
		load_wb(X.B, r.z, b);
		m3d_vec3 pvA = f4_v3(X.B.prevVel[r.y]), pwA = f4_v3(X.B.prevAngVel[r.y]);
		m3d_vec3 pvB = f4_v3(X.B.prevVel[r.z]), pwB = f4_v3(X.B.prevAngVel[r.z]);
		// issued with the other body loads (one round trip), although only
		// curved pairs use them
		VelShape sa, sbv;
		sa.type = bf_shape(X.B.flags[r.y]);
		sa.radius = X.B.shape[r.y].w;
		sa.halfHeight = X.B.mat[r.y].x;
		sbv.type = bf_shape(X.B.flags[r.z]);
		sbv.radius = X.B.shape[r.z].w;
		sbv.halfHeight = X.B.mat[r.z].x;
		solve_velocity_in(X.C, r.x, d, X.h, X.gravityLen, a, b, pvA, pwA, pvB, pwB, sa, sbv);
		store_vel(X.B, r.y, a);
		store_vel(X.B, r.z, b);
	}
};

template <class Pass>
__device__ __forceinline__ void run_rec(const PassCtx& X, int4 r)
{
	typename Pass::In d;
	Pass::load(X, r, d);
	Pass::run(X, r, d);
}

__device__ void pass_normals_serial(const PassCtx& X, int32_t c) { run_rec<NormalsPass>(X, make_rec(X, c, -1)); }
__device__ void pass_normals_packet(const PassCtx& X, int32_t c, int32_t maxPoints)
{
	run_rec<NormalsPass>(X, make_rec(X, c, maxPoints));
}
__device__ void pass_friction(const PassCtx& X, int32_t c) { run_rec<FrictionPass>(X, make_rec(X, c, -1)); }
__device__ void pass_velocity(const PassCtx& X, int32_t c) { run_rec<VelocityPass>(X, make_rec(X, c, -1)); }

__device__ void pass_joint(const PassCtx& X, DJoint& jt, int op)
{
	WB a, b;
	load_wb(X.B, jt.bodyA, a);
	load_wb(X.B, jt.bodyB, b);
	if (op == 0)
	{
		prepare_joint(jt, a, b);
	}
	else if (op == 1)
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

struct IslandCtx
{
	PassCtx P;
	const uint32_t* bodyList;
	const uint32_t* bodyStart;
	const uint32_t* contactList;
	const uint32_t* contactStart;
	const uint64_t* entryIsland;
	const uint32_t* mode;
	const int32_t* islandOf;
	DJoint* J;
	int32_t nJ;
	const int32_t* jointIsland;
	int32_t* contactColor;
	m3d_vec3 gravity;
	float invH;
	int32_t iterations, substeps, I, NB, NC;
	uint32_t xcdBar;     // grid barriers in XCD mode (GridSync)
	uint32_t barTimeout; // longest grid-barrier wait, in wall-clock ticks (GridSync)
};

// ------------------------------------------------------------- joints

// Bodies of awake islands with joints: world inverse inertia refreshed once
// per step before the joints are prepared (src/xpbd.cpp:909-921).
__global__ void k_joint_island_inertia(IslandCtx X)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= X.NB)
		return;
	int32_t b = (int32_t)X.bodyList[k];
	int32_t isl = X.islandOf[b];
	if (isl >= 0 && (X.mode[isl] & (IM_AWAKE | IM_JOINT)) == (IM_AWAKE | IM_JOINT))
		store_iw(X.P.B, b, inv_inertia_world(f4_q(X.P.B.rot[b]), f4_v3(X.P.B.invIL[b])));
}

// op 0: PrepareJoint of every joint of an awake island (each reads only its
// own bodies); op 3: the substep's impulse reset.
__global__ void k_joints_all(IslandCtx X, int op)
{
	int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (j >= X.nJ || X.jointIsland[j] < 0)
		return;
	if (op == 0)
	{
		pass_joint(X.P, X.J[j], 0);
	}
	else
	{
		X.J[j].lambda = 0.0f;
		X.J[j].motorLambda = 0.0f;
	}
}

// One level of the joints' serial order: op 1 the position solve, 2 the
// velocity solve.
__global__ void k_joint_level(IslandCtx X, const int32_t* order, int32_t begin, int32_t end, int op)
{
	int32_t k = begin + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= end)
		return;
	int32_t j = order[k];
	if (X.jointIsland[j] < 0)
		return;
	pass_joint(X.P, X.J[j], op);
}

// ------------------------------------------------------ giant islands

constexpr int kGiantThreads = 256;

struct GiantCtx
{
	IslandCtx X;
	const uint32_t* adjStart; // per body (n + 1)
	const uint32_t* adj;      // contact-list entries sorted by (body, entry)
	const int4* pred;         // per entry: predecessor ranges in adj (x < 0: not on the grid path)
	const uint8_t* colored;   // per entry: island takes greedy colors
	int32_t* color;           // per entry
	int32_t* level;           // per entry
	uint32_t* ctr;            // 6 convergence counters (zeroed before launch)
	uint32_t* bar;            // GridSync counters (kBarWords, zeroed before launch)
	uint32_t* counters;
	uint64_t* slotKeys;       // per entry, sorted later
	uint32_t* slotVals;
	uint8_t* dirty;           // 2 x NC: entries to re-evaluate in the next round (zeroed)
	int bitsI;
	int32_t maxS;             // slot bound: slots are 0 .. maxS - 1
};

__device__ __forceinline__ bool giant_entry(const GiantCtx& G, int32_t k)
{
	return (G.X.mode[(int32_t)G.X.entryIsland[k]] & IM_GRID) != 0;
}

// Barrier of a co-resident grid, in two levels: workgroup b arrives on the
// counter of group b % kBarGroups, the last arrival of a group on the top
// counter, and every workgroup waits for the top counter. Each counter has
// its own 128-byte line: device-scope atomics on one address from every
// workgroup serialize (about 0.04 us each), and the two levels leave at most
// blocks / kBarGroups + kBarGroups of them in a row. All counters are
// monotonic (zeroed before the launch): barrier e completes when group g
// reaches (e + 1) * size(g) and the top counter (e + 1) * groups. The
// release fence publishes this workgroup's writes before it arrives, the
// group's last arrival passes them on (acquire-release) to the top counter,
// and the acquire fence drops stale cache lines before a workgroup reads
// others' writes.
//
// XCD mode (xcd != 0). The MI355X's 8 XCDs each have their own L2. An
// agent-scope release writes the whole L2 back (buffer_wbl2) and an acquire
// invalidates it (buffer_inv sc1); with every workgroup doing both, a barrier
// cost 3.3 us at 32 workgroups and 10.3 us at 256 (the write-backs serialize
// in each L2). In XCD mode the groups are the XCDs themselves: at the launch
// every workgroup reads its XCD (the XCC_ID register) and counts itself in a
// census, one barrier of the plain kind makes the census final, and a group's
// size is its census count (the dispatcher's XCD for workgroup b rotates from
// launch to launch, so b % 8 is not the XCD). Every wave waits for its own
// stores to reach the L2 before the workgroup arrives, and only the group's
// last arrival writes the L2 back (the group's stores with it) before
// arriving on the top counter: one write-back per XCD. Every workgroup still
// invalidates for itself after the top counter: an invalidation orders only
// the issuing CU's later loads, and a variant with one invalidation per XCD
// handed on by a flag let other CUs read stale lines (a coherence probe with
// concurrent readers counted them; none with per-workgroup invalidation).
// Measured: 3.0 us at 64 workgroups (4.3 plain), 4.7 us at 256 (10.3).
// Every group counter receives exactly its census count of arrivals per
// barrier and the top counter one per nonempty group, so the waits end as in
// the plain kind.
constexpr uint32_t kBarGroups = 8;
constexpr uint32_t kBarStride = 32; // u32 per counter
// counters (kBarStride words apart): group arrivals, top, leave, the XCD
// census, and the census barrier's group arrivals and top
constexpr uint32_t kBarTop = kBarGroups;
constexpr uint32_t kBarLeave = kBarGroups + 1;
constexpr uint32_t kBarCensus0 = kBarGroups + 2;
constexpr uint32_t kBarCGroup0 = 2 * kBarGroups + 2;
constexpr uint32_t kBarCTop = 3 * kBarGroups + 2;
constexpr uint32_t kBarCounters = 3 * kBarGroups + 3;
constexpr uint32_t kBarWords = kBarCounters * kBarStride;

__device__ __forceinline__ uint32_t xcc_id()
{
	// hwreg(HW_REG_XCC_ID, 0, 4)
	return __builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11));
}

// Bounded waits. The grid's workgroups are co-resident by its size (at most
// what the occupancy of every CU holds), so every wait ends; should one not
// end within `limit` wall-clock ticks (s_memrealtime, 100 MHz), the waiting
// lane sets the step's abort word and every later wait of the step returns
// at once: the launch drains and the step reports the failure. Returns
// false when the wait was cut short.
__device__ __forceinline__ bool bounded_wait(const uint32_t* word, uint32_t target, uint32_t* abortw, uint32_t limit)
{
	uint64_t t0 = 0;
	for (uint32_t k = 0;; ++k)
	{
		uint32_t v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		uint32_t a = __hip_atomic_load(abortw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (v >= target)
			return true;
		if (a != 0u)
			return false;
		uint64_t t = __builtin_amdgcn_s_memrealtime();
		if (k == 0)
			t0 = t;
		else if (t - t0 > (uint64_t)limit)
		{
			__hip_atomic_store(abortw, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			return false;
		}
		__builtin_amdgcn_s_sleep(1);
	}
}

// Lane 0 of every workgroup: one two-level barrier of the plain kind over
// groups b % kBarGroups on the counters at groupBase / top (epoch `e`).
__device__ __forceinline__ void plain_barrier_lane0(uint32_t* groupBase, uint32_t* top, uint32_t e, uint32_t* abortw,
													uint32_t limit)
{
	const uint32_t blocks = gridDim.x;
	const uint32_t g = blockIdx.x % kBarGroups;
	const uint32_t size = blocks / kBarGroups + (g < blocks % kBarGroups ? 1u : 0u);
	const uint32_t groups = blocks < kBarGroups ? blocks : kBarGroups;
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	uint32_t old = __hip_atomic_fetch_add(groupBase + g * kBarStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	if (old + 1u == (e + 1u) * size)
	{
		__builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
		__hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
	(void)bounded_wait(top, (e + 1u) * groups, abortw, limit);
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

struct GridSync
{
	uint32_t* base; // kBarWords words, zero at the launch
	uint32_t epoch;
	uint32_t xcd;    // XCD mode
	uint32_t grp;    // this workgroup's group
	uint32_t size;   // workgroups in the group
	uint32_t groups; // nonempty groups
	uint32_t vblock; // this workgroup's place in the launch's XCD-major order (XCD mode), else blockIdx.x
	uint32_t* abortw; // the step's abort word (bounded_wait)
	uint32_t limit;   // wait bound, wall-clock ticks

	// Called by every thread of the workgroup at the launch.
	__device__ static GridSync make(uint32_t* base, uint32_t xcdMode, uint32_t* abortw, uint32_t limit)
	{
		GridSync g;
		g.base = base;
		g.epoch = 0u;
		g.xcd = xcdMode;
		g.abortw = abortw;
		g.limit = limit;
		const uint32_t blocks = gridDim.x;
		if (xcdMode == 0u)
		{
			g.grp = blockIdx.x % kBarGroups;
			g.size = blocks / kBarGroups + (g.grp < blocks % kBarGroups ? 1u : 0u);
			g.groups = blocks < kBarGroups ? blocks : kBarGroups;
			g.vblock = blockIdx.x;
			return g;
		}
		__shared__ uint32_t census[4];
		if (threadIdx.x == 0)
		{
			const uint32_t x = xcc_id() % kBarGroups;
			const uint32_t rank = __hip_atomic_fetch_add(base + (kBarCensus0 + x) * kBarStride, 1u, __ATOMIC_RELAXED,
														 __HIP_MEMORY_SCOPE_AGENT);
			plain_barrier_lane0(base + kBarCGroup0 * kBarStride, base + kBarCTop * kBarStride, 0u, abortw, limit);
			uint32_t nonempty = 0u, mine = 0u, before = 0u;
			for (uint32_t k = 0; k < kBarGroups; ++k)
			{
				uint32_t c = __hip_atomic_load(base + (kBarCensus0 + k) * kBarStride, __ATOMIC_RELAXED,
											   __HIP_MEMORY_SCOPE_AGENT);
				nonempty += c != 0u ? 1u : 0u;
				mine = k == x ? c : mine;
				before += k < x ? c : 0u;
			}
			census[0] = x;
			census[1] = mine;
			census[2] = nonempty;
			census[3] = before + rank;
		}
		__syncthreads();
		g.grp = census[0];
		g.size = census[1];
		g.groups = census[2];
		g.vblock = census[3];
		return g;
	}

	__device__ uint32_t* top() const { return base + kBarTop * kBarStride; }

	__device__ void arrive_lane0()
	{
		if (xcd == 0u)
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
		uint32_t old = __hip_atomic_fetch_add(base + grp * kBarStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (old + 1u == (epoch + 1u) * size)
		{
			if (xcd != 0u)
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
			else
				__builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
			__hip_atomic_fetch_add(top(), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
	}

	__device__ bool wait_lane0() const
	{
		bool ok = bounded_wait(top(), (epoch + 1u) * groups, abortw, limit);
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
		return ok;
	}

	// every wave: its stores have reached the L2 (XCD mode; the agent release
	// of the plain kind covers the workgroup)
	__device__ void stores_done() const
	{
		if (xcd != 0u)
			__builtin_amdgcn_s_waitcnt(0);
	}

	// Barrier; then returns *word as read by one lane after it (the same
	// value for every lane of the grid as long as nobody writes *word
	// before the next barrier), or 0 when the wait was cut short.
	__device__ uint32_t sync_read(const uint32_t* word)
	{
		__shared__ uint32_t value;
		stores_done();
		__syncthreads();
		if (threadIdx.x == 0)
		{
			arrive_lane0();
			bool ok = wait_lane0();
			if (word)
				value = ok ? __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
		}
		++epoch;
		__syncthreads();
		return word ? value : 0u;
	}

	__device__ void sync() { (void)sync_read(nullptr); }

	// The same barrier in two halves: loads that do not depend on other
	// workgroups' writes of this phase may be issued in between.
	__device__ void arrive()
	{
		stores_done();
		__syncthreads();
		if (threadIdx.x == 0)
			arrive_lane0();
	}

	__device__ void wait()
	{
		if (threadIdx.x == 0)
			(void)wait_lane0();
		++epoch;
		__syncthreads();
	}

	// the last workgroup to leave the launch zeroes the counters for the next
	__device__ void leave() const
	{
		if (threadIdx.x != 0)
			return;
		uint32_t left =
			__hip_atomic_fetch_add(base + kBarLeave * kBarStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (left + 1 == gridDim.x)
			for (uint32_t k = 0; k < kBarCounters; ++k)
				__hip_atomic_store(base + k * kBarStride, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
};

// Plain loads and stores of values that other lanes update within a
// fixpoint round: a round may see either value, and the grid barrier
// between rounds publishes every update.
__device__ __forceinline__ int32_t ld_round(const int32_t* p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void st_round(int32_t* p, int32_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// OR of the colors (< 8) of the entries adj[s..e): the entry's predecessors
// on one body. Loads are issued 8 at a time.
__device__ __forceinline__ uint32_t pred_colors(const GiantCtx& G, int32_t s, int32_t e)
{
	uint32_t used = 0;
	for (int32_t base = s; base < e; base += 8)
	{
		int32_t k2[8];
#pragma unroll
		for (int i = 0; i < 8; ++i)
			k2[i] = base + i < e ? (int32_t)G.adj[base + i] : -1;
#pragma unroll
		for (int i = 0; i < 8; ++i)
		{
			if (k2[i] >= 0)
			{
				int32_t col = ld_round(&G.color[k2[i]]);
				if (col < kMaxGraphColors)
					used |= 1u << col;
			}
		}
	}
	return used;
}

// Deepest overflow level + 1 among the predecessors adj[s..e).
__device__ __forceinline__ int32_t pred_level(const GiantCtx& G, int32_t s, int32_t e)
{
	int32_t lv = 0;
	for (int32_t base = s; base < e; base += 8)
	{
		int32_t k2[8];
#pragma unroll
		for (int i = 0; i < 8; ++i)
			k2[i] = base + i < e ? (int32_t)G.adj[base + i] : -1;
#pragma unroll
		for (int i = 0; i < 8; ++i)
		{
			if (k2[i] >= 0 && G.color[k2[i]] == kMaxGraphColors)
			{
				int32_t l2 = ld_round(&G.level[k2[i]]) + 1;
				lv = l2 > lv ? l2 : lv;
			}
		}
	}
	return lv;
}

// One coloring round for entry k: first fit over its predecessors (earlier
// entries sharing a written body, the ranges pr); true when it changed.
__device__ bool color_round(const GiantCtx& G, int32_t k, int4 pr)
{
	uint32_t used = pred_colors(G, pr.x, pr.y) | pred_colors(G, pr.z, pr.w);
	int32_t nc = (int32_t)__builtin_ctz(~used);
	nc = nc < kMaxGraphColors ? nc : kMaxGraphColors;
	if (nc != ld_round(&G.color[k]))
	{
		st_round(&G.color[k], nc);
		return true;
	}
	return false;
}

// One level round for an overflow entry (serial order among the overflow).
__device__ bool level_round(const GiantCtx& G, int32_t k, int4 pr)
{
	int32_t la = pred_level(G, pr.x, pr.y), lb = pred_level(G, pr.z, pr.w);
	int32_t lv = la > lb ? la : lb;
	if (lv != ld_round(&G.level[k]))
	{
		st_round(&G.level[k], lv);
		return true;
	}
	return false;
}

// An entry changed: every later entry on its dynamic bodies (its successors,
// which read it) is re-evaluated in the next round.
__device__ void mark_successors(const GiantCtx& G, int32_t k, int4 pr, uint8_t* next)
{
	int2 ab = G.X.P.C.bodies[G.X.contactList[k]];
	if (G.X.P.B.pos[ab.x].w > 0.0f)
	{
		int32_t end = (int32_t)G.adjStart[ab.x + 1];
		for (int32_t e = pr.y + 1; e < end; ++e)
			next[G.adj[e]] = 1u;
	}
	if (G.X.P.B.pos[ab.y].w > 0.0f)
	{
		int32_t end = (int32_t)G.adjStart[ab.y + 1];
		for (int32_t e = pr.w + 1; e < end; ++e)
			next[G.adj[e]] = 1u;
	}
}

// Rounds until one changes nothing. Round 0 evaluates every entry; round
// r > 0 only the entries one of whose predecessors changed in round r - 1
// (dirty[r % 2], cleared as they are taken; marks for the next round go to
// the other half). Round r flags changes in ctr[r % 3] and clears
// ctr[(r + 1) % 3], which no lane reads again before it is reused two
// rounds later.
// The entries taken are [k0, k1) (all of them on the grid, one sub-world's
// on a workgroup: no predecessor or successor of an entry lies in another
// sub-world), tid and nt the lane's place and count over them; Sync is the
// grid barrier or the workgroup's (BlockSync).
template <int WHICH, class Sync>
__device__ int32_t fixpoint_rounds(const GiantCtx& G, Sync& grid, uint32_t* ctr, int32_t tid, int32_t nt, int32_t k0,
								   int32_t k1)
{
	const int32_t NC = G.X.NC;
	for (int32_t r = 0; r < (1 << 24); ++r)
	{
		if (tid == 0)
			__hip_atomic_store(&ctr[(r + 1) % 3], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		uint8_t* cur = G.dirty + (size_t)(r & 1) * NC;
		uint8_t* next = G.dirty + (size_t)((r + 1) & 1) * NC;
		uint32_t changes = 0;
		for (int32_t k = k0 + tid; k < k1; k += nt)
		{
			if (r > 0)
			{
				if (!cur[k])
					continue;
				cur[k] = 0u;
			}
			int4 pr = G.pred[k];
			if (pr.x < 0)
				continue;
			bool changed = false;
			if (WHICH == 0)
			{
				if (G.colored[k])
					changed = color_round(G, k, pr);
			}
			else if (G.color[k] == kMaxGraphColors)
				changed = level_round(G, k, pr);
			if (changed)
			{
				++changes;
				mark_successors(G, k, pr, next);
			}
		}
		// one flag store per workgroup that changed something (same-address
		// atomics from every wave would serialize)
		if (__syncthreads_or(changes != 0u) && threadIdx.x == 0)
			__hip_atomic_store(&ctr[r % 3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (grid.sync_read(&ctr[r % 3]) == 0u)
			return r + 1;
	}
	return -1;
}

// Slot keys, (slot, island) for giant entries and a sentinel above all
// others, of the entries [k0, k1); returns the lane's largest slot + 1.
__device__ uint32_t order_keys(const GiantCtx& G, int32_t tid, int32_t nt, int32_t k0, int32_t k1)
{
	uint32_t maxSlot = 0;
	for (int32_t k = k0 + tid; k < k1; k += nt)
	{
		uint64_t key = ~0ull;
		if (giant_entry(G, k))
		{
			int32_t col = G.color[k];
			int32_t slot = col < kMaxGraphColors ? col : kMaxGraphColors + G.level[k];
			// within the slot arrays whatever the rounds left (a fixpoint
			// cut short by an aborted barrier); complete rounds stay below
			slot = slot < 0 ? 0 : (slot >= G.maxS ? G.maxS - 1 : slot);
			if (G.colored[k])
				G.X.contactColor[G.X.contactList[k]] = col;
			G.X.P.C.hint[G.X.contactList[k]] = make_int2(col, G.level[k]);
			key = ((uint64_t)slot << G.bitsI) | G.X.entryIsland[k];
			maxSlot = (uint32_t)slot + 1u > maxSlot ? (uint32_t)slot + 1u : maxSlot;
		}
		G.slotKeys[k] = key;
		G.slotVals[k] = G.X.contactList[k];
	}
	return maxSlot;
}

__global__ void __launch_bounds__(kGiantThreads) k_giant_order(GiantCtx G)
{
	GridSync grid = GridSync::make(G.bar, G.X.xcdBar, G.counters + CT_GRID_ABORT, G.X.barTimeout);
	int32_t tid = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	int32_t nt = (int32_t)(gridDim.x * blockDim.x);
	int32_t rc = fixpoint_rounds<0>(G, grid, G.ctr, tid, nt, 0, G.X.NC);
	int32_t rl = fixpoint_rounds<1>(G, grid, G.ctr + 3, tid, nt, 0, G.X.NC);
	if (tid == 0)
		G.counters[CT_COLOR_ROUNDS] = (uint32_t)(rc + rl);
	uint32_t maxSlot = order_keys(G, tid, nt, 0, G.X.NC);
	// one atomic per wave (all lanes reach here)
	for (int off = 32; off > 0; off >>= 1)
	{
		uint32_t o = __shfl_xor(maxSlot, off);
		maxSlot = o > maxSlot ? o : maxSlot;
	}
	if (__lane_id() == 0 && maxSlot > 0u)
		atomicMax(&G.counters[CT_SLOTS], maxSlot);
}

// Batched worlds: a sub-world's islands are contiguous island ids and their
// entries a contiguous range, so one workgroup per sub-world runs the same
// fixpoints over its range with workgroup barriers (the grid barriers of
// k_giant_order cost microseconds each, over rounds of a whole batch).
struct BlockSync
{
	__device__ uint32_t sync_read(const uint32_t* word)
	{
		__syncthreads();
		uint32_t v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		__syncthreads();
		return v;
	}
};

// start[s] = first entry of sub-world s or later (entries ascend by
// sub-world), start[S] = NC.
__global__ void k_sub_entry_start(const GiantCtx G, const int32_t* bodySub, int32_t S, int32_t* start)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	const int32_t NC = G.X.NC;
	if (k > NC)
		return;
	int32_t hi = k < NC ? bodySub[G.X.P.C.bodies[G.X.contactList[k]].x] : S;
	int32_t lo = k > 0 ? bodySub[G.X.P.C.bodies[G.X.contactList[k - 1]].x] : -1;
	for (int32_t sw = lo + 1; sw <= hi; ++sw)
		start[sw] = k;
}

__global__ void __launch_bounds__(kGiantThreads) k_giant_order_sub(GiantCtx G, const int32_t* subEntryStart)
{
	__shared__ uint32_t ctr[6];
	const int32_t k0 = subEntryStart[blockIdx.x], k1 = subEntryStart[blockIdx.x + 1];
	if (threadIdx.x < 6u)
		ctr[threadIdx.x] = 0u;
	__syncthreads();
	BlockSync sync;
	const int32_t tid = (int32_t)threadIdx.x, nt = (int32_t)blockDim.x;
	int32_t rc = fixpoint_rounds<0>(G, sync, ctr, tid, nt, k0, k1);
	int32_t rl = fixpoint_rounds<1>(G, sync, ctr + 3, tid, nt, k0, k1);
	uint32_t maxSlot = order_keys(G, tid, nt, k0, k1);
	maxSlot = wave_max(maxSlot);
	__shared__ uint32_t wmax[kGiantThreads / 64];
	if (__lane_id() == 0)
		wmax[threadIdx.x >> 6] = maxSlot;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		for (uint32_t w = 1; w < (blockDim.x + 63) / 64; ++w)
			maxSlot = wmax[w] > maxSlot ? wmax[w] : maxSlot;
		if (maxSlot > 0u)
			atomicMax(&G.counters[CT_SLOTS], maxSlot);
		// rounds of the sub-world that took the most
		if (k1 > k0)
			atomicMax(&G.counters[CT_COLOR_ROUNDS], (uint32_t)(rc + rl));
	}
}

// Adjacency pairs (body, entry) of the giant entries, and the fixpoint's
// starting point: the contact's color and level of its last solve (any start
// converges to the same least fixpoint; a close one in fewer rounds).
__global__ void k_giant_adj(IslandCtx X, int4* pred, uint8_t* colored, int32_t* color, int32_t* level, int32_t n,
							uint64_t* keys, uint32_t* vals, uint32_t* bodyCount)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= X.NC)
		return;
	uint32_t m = X.mode[(int32_t)X.entryIsland[k]];
	bool g = (m & IM_GRID) != 0;
	bool col = g && (m & IM_COLORED);
	int2 ab = X.P.C.bodies[X.contactList[k]];
	bool da = g && X.P.B.pos[ab.x].w > 0.0f;
	bool db = g && X.P.B.pos[ab.y].w > 0.0f;
	keys[2 * k] = da ? (uint64_t)ab.x : (uint64_t)n;
	vals[2 * k] = (uint32_t)k;
	keys[2 * k + 1] = db ? (uint64_t)ab.y : (uint64_t)n;
	vals[2 * k + 1] = (uint32_t)k;
	if (da)
		atomicAdd(&bodyCount[ab.x], 1u);
	if (db)
		atomicAdd(&bodyCount[ab.y], 1u);
	pred[k] = g ? make_int4(0, 0, 0, 0) : make_int4(-1, -1, -1, -1);
	colored[k] = col ? 1u : 0u;
	int2 h = g ? X.P.C.hint[X.contactList[k]] : make_int2(0, 0);
	// islands below the colored threshold: every contact in the serial pass
	color[k] = !col ? kMaxGraphColors : (h.x < 0 ? 0 : (h.x > kMaxGraphColors ? kMaxGraphColors : h.x));
	level[k] = h.y < 0 ? 0 : (h.y > (1 << 24) ? (1 << 24) : h.y);
}

// Predecessor range of every (body, entry) pair: the entries before it in
// the body's sorted list.
__global__ void k_giant_pred(IslandCtx X, const uint64_t* keys, const uint32_t* adj, const uint32_t* adjStart,
							 int64_t pairs, int32_t n, int4* pred)
{
	int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= pairs || keys[e] >= (uint64_t)n)
		return;
	int32_t body = (int32_t)keys[e];
	int32_t k = (int32_t)adj[e];
	int32_t st = (int32_t)adjStart[body];
	if (X.P.C.bodies[X.contactList[k]].x == body)
	{
		pred[k].x = st;
		pred[k].y = (int32_t)e;
	}
	else
	{
		pred[k].z = st;
		pred[k].w = (int32_t)e;
	}
}

// slotStart[s] = first sorted entry with slot >= s, for s in [0, maxS]
__global__ void k_giant_slot_starts(const uint64_t* keys, int32_t NC, int bitsI, int32_t maxS, uint32_t* slotStart)
{
	int32_t s = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (s > maxS)
		return;
	uint64_t target = (uint64_t)s << bitsI;
	int32_t l = 0, r = NC;
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

// packet = 8 consecutive members of one (color, island) bucket; -1 outside
// the colored slots
__global__ void k_giant_packets(const uint64_t* keys, int32_t NC, int bitsI, int2* packet)
{
	int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (e >= NC)
		return;
	uint64_t key = keys[e];
	if (key == ~0ull || (int32_t)(key >> bitsI) >= kMaxGraphColors)
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
	r = NC;
	while (l < r)
	{
		int32_t mid = (l + r) >> 1;
		if (keys[mid] <= key)
			l = mid + 1;
		else
			r = mid;
	}
	int32_t ps = gs + ((e - gs) & ~7);
	int32_t pe = ps + 8 < l ? ps + 8 : l;
	packet[e] = make_int2(ps, pe);
}

struct GiantSolveCtx
{
	IslandCtx X;
	const uint32_t* entries;   // contact per sorted entry
	const int2* packet;
	const uint32_t* slotStart; // [0, nSlots]
	const uint32_t* counters;
	uint32_t* counters2;       // the same block, for the record counters
	uint32_t* barrier;         // GridSync counters (kBarWords), zero between launches
	int4* rec;                 // solve records, grouped within each slot
	const uint64_t* keys;      // sorted (slot, island) keys
	int bitsI;
	uint32_t* bucketCount;     // per (slot, group); zero between steps
	uint32_t* bucketCursor;    // per (slot, group)
	uint32_t* pos;             // record position of each sorted entry (per step)
	uint32_t* cnt;             // point count of each sorted entry (per substep)
	uint32_t* entryAt;         // sorted entry of each record position (per step)
	uint32_t* atOf;            // record position of each grid contact, by contact (per step)
	// contact-side inputs of the three passes for each scheduled contact at
	// its record position (block b of record k at b * nrmStride + k):
	// 0 normal, 1-4 anchors A (+ separation offset), 5-8 anchors B (+ wN),
	// 9-12 angJA (+ lambdaN), 13-16 angJB (+ lambdaT), 17 material, 18-21
	// slip directions, 22-25 / 26-29 frozen arms A / B (flat contacts),
	// 30-33 / 34-37 friction anchors A / B (box-box contacts). The passes
	// read them coalesced; the normal pass accumulates lambdaN there and
	// k_giant_lambda_back returns it to the contacts. Everything a pass
	// writes besides lambdaN (lambdaT, re-anchoring, slip) goes to the
	// contacts, which only the contact itself reads later in the substep.
	float4* nrm;
	int64_t nrmStride;
	// the substep returns lambdaN from the copy to the contacts (the last
	// one only: each substep's prepare zeroes it again, and within a
	// substep every pass reads it from the copy)
	int lambdaBack;
};

constexpr int kSchedBlocks = 38;

// Contact-side inputs of a scheduled record e (its contact r.x) and the pass
// itself; the normal pass reads the schedule-ordered copy.
template <class Pass>
__device__ __forceinline__ void sweep_load(const GiantSolveCtx& S, int32_t e, int4 r, typename Pass::In& d)
{
	Pass::load(S.X.P, r, d);
}

template <class Pass>
__device__ __forceinline__ void sweep_run(const GiantSolveCtx& S, int32_t e, int4 r, const typename Pass::In& d)
{
	Pass::run(S.X.P, r, d);
}

// Only the contact's own points are loaded (the copy holds nothing beyond
// them, and the passes read nothing beyond them); records of one (slot, point
// count) bucket are adjacent, so whole waves skip the same columns.
template <>
__device__ __forceinline__ void sweep_load<NormalsPass>(const GiantSolveCtx& S, int32_t e, int4 r, NormalsIn& d)
{
	const int64_t st = S.nrmStride;
	const int n = rec_points(r);
	const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	d.nc = S.nrm[e];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		bool on = j < n;
		d.A4[j] = on ? S.nrm[(1 + j) * st + e] : zero;
		d.B4[j] = on ? S.nrm[(5 + j) * st + e] : zero;
		d.JA[j] = on ? S.nrm[(9 + j) * st + e] : zero;
		d.JB[j] = on ? S.nrm[(13 + j) * st + e] : zero;
	}
}

template <>
__device__ __forceinline__ void sweep_load<FrictionPass>(const GiantSolveCtx& S, int32_t e, int4 r, FrictionIn& d)
{
	if (!(r.w & kRecBoxBox))
		return;
	const int64_t st = S.nrmStride;
	const int n = rec_points(r);
	const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	d.nc = S.nrm[e];
	d.mu = S.nrm[17 * st + e].x;
	d.RA0 = S.nrm[22 * st + e];
	d.RB0 = S.nrm[26 * st + e];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		bool on = j < n;
		d.LN[j] = on ? S.nrm[(9 + j) * st + e].w : 0.0f;
		d.FA[j] = on ? S.nrm[(30 + j) * st + e] : zero;
		d.FB[j] = on ? S.nrm[(34 + j) * st + e] : zero;
		d.JB[j] = on ? S.nrm[(13 + j) * st + e] : zero;
	}
}

// All four points' columns, unguarded: a point-count guard here made the
// pass slower (it outweighed the skipped columns). The curved bit comes from
// the record, so no load waits on another load of this contact.
template <>
__device__ __forceinline__ void sweep_load<VelocityPass>(const GiantSolveCtx& S, int32_t e, int4 r, VelocityIn& d)
{
	const int64_t st = S.nrmStride;
	d.nc = S.nrm[e];
	d.mat = S.nrm[17 * st + e];
	bool curved = (r.w & kRecCurved) != 0;
	const int64_t ra = curved ? 1 : 22, rb = curved ? 5 : 26;
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		d.JA[j] = S.nrm[(9 + j) * st + e];
		d.B4[j] = S.nrm[(5 + j) * st + e];
		d.SL[j] = S.nrm[(18 + j) * st + e];
		d.JBv[j] = S.nrm[(13 + j) * st + e];
		d.AR[j] = S.nrm[(ra + j) * st + e];
		d.BR[j] = S.nrm[(rb + j) * st + e];
	}
}

template <>
__device__ __forceinline__ void sweep_run<NormalsPass>(const GiantSolveCtx& S, int32_t e, int4 r, const NormalsIn& d)
{
	// lambdaN accumulates in the copy: a contact view whose angJA column is
	// the copy's block 9, indexed by record position
	PassCtx P = S.X.P;
	P.C.pJA = S.nrm + 9 * S.nrmStride;
	P.C.cap = (int32_t)S.nrmStride;
	P.C.cs = 1;
	P.C.js = (int32_t)S.nrmStride;
	NormalsPass::run(P, make_int4(e, r.y, r.z, r.w), d);
}

// Slots with at most this many entries are run by workgroup 0 alone: a run
// of consecutive small slots then costs one grid barrier instead of one per
// slot (the overflow levels are many and small).
constexpr int32_t kSmallSlot = 256;

// First entry of slot s for this lane, or -1: every lane of the grid for
// large slots, the lanes of the first workgroup for small ones. ss: slot
// starts; tid: the lane's place in the launch's XCD-major order (GridSync
// vblock), so consecutive records of a slot go to workgroups of one XCD and
// the body and record lines they share are fetched into one L2.
__device__ __forceinline__ int32_t giant_lane_entry(const uint32_t* ss, int32_t s, int32_t tid, int32_t& stride)
{
	int32_t st = (int32_t)ss[s], en = (int32_t)ss[s + 1];
	int32_t e;
	if (en - st <= kSmallSlot)
	{
		if (tid >= (int32_t)blockDim.x)
			return -1;
		e = st + (int32_t)threadIdx.x;
		stride = (int32_t)blockDim.x;
	}
	else
	{
		e = st + tid;
		stride = (int32_t)(gridDim.x * blockDim.x);
	}
	return e < en ? e : -1;
}

// One Gauss-Seidel pass over all slots. Each lane's first entry of the next
// slot (its record and contact-side inputs, which no other contact writes)
// is loaded while the grid barrier that ends the current slot is pending.
template <class Pass>
__device__ void giant_sweep(const GiantSolveCtx& S, const uint32_t* ss, GridSync& grid, int32_t nSlots)
{
	const PassCtx& P = S.X.P;
	const int32_t tid = (int32_t)(grid.vblock * blockDim.x + threadIdx.x);
	int32_t stride = 1;
	int4 r = make_int4(0, 0, 0, 0);
	typename Pass::In d;
	int32_t e = nSlots > 0 ? giant_lane_entry(ss, 0, tid, stride) : -1;
	if (e >= 0)
	{
		r = S.rec[e];
		sweep_load<Pass>(S, e, r, d);
	}
	bool pendingSmall = false;
	for (int32_t s = 0; s < nSlots; ++s)
	{
		int32_t st = (int32_t)ss[s], en = (int32_t)ss[s + 1];
		bool small = en - st <= kSmallSlot;
		if (st != en && !small && pendingSmall)
		{
			grid.sync();
			pendingSmall = false;
		}
		if (e >= 0)
		{
			sweep_run<Pass>(S, e, r, d);
			for (int32_t e2 = e + stride; e2 < en; e2 += stride)
			{
				typename Pass::In d2;
				int4 r2 = S.rec[e2];
				sweep_load<Pass>(S, e2, r2, d2);
				sweep_run<Pass>(S, e2, r2, d2);
			}
		}
		int32_t strideNext = stride;
		int32_t eNext = s + 1 < nSlots ? giant_lane_entry(ss, s + 1, tid, strideNext) : -1;
		bool big = st != en && !small;
		if (big)
			grid.arrive();
		else if (st != en && grid.vblock == 0)
			__syncthreads();
		e = eNext;
		stride = strideNext;
		if (e >= 0)
		{
			r = S.rec[e];
			sweep_load<Pass>(S, e, r, d);
		}
		if (big)
			grid.wait();
		else if (st != en)
			pendingSmall = true;
	}
	if (pendingSmall)
		grid.sync();
}

// Substep stages of the giant islands: plain launches for the per-body and
// per-contact stages, one cooperative launch per Gauss-Seidel pass (so each
// holds only that pass's registers).
__device__ __forceinline__ bool giant_body(const IslandCtx& X, int32_t b)
{
	return (X.mode[X.islandOf[b]] & IM_GRID) != 0;
}

__global__ void k_giant_bodies(GiantSolveCtx S, int op)
{
	const IslandCtx& X = S.X;
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= X.NB)
		return;
	int32_t b = (int32_t)X.bodyList[k];
	if (!giant_body(X, b))
		return;
	if (op == 0)
		integrate_body(X.P.B, b, X.P.h, X.gravity);
	else
		reconstruct_velocity(X.P.B, b, X.invH);
}

// Within a slot the contacts are independent, so their order is free: the
// sweeps run them grouped by point count and curvature (kBuckets groups per
// slot), which keeps the lanes of a wave on the same path.
constexpr int kBuckets = 8;

__device__ __forceinline__ uint32_t giant_bucket(const PassCtx& P, const GiantSolveCtx& S, int32_t e, int32_t c)
{
	int32_t pc = f2i(P.C.normal[c].w);
	pc = pc < 1 ? 1 : (pc > 4 ? 4 : pc);
	bool curved = (f2u(P.C.mat[c].w) & CF_CURVED) != 0;
	uint32_t slot = (uint32_t)(S.keys[e] >> S.bitsI);
	return slot * kBuckets + (uint32_t)(pc - 1) * 2u + (curved ? 1u : 0u);
}

// atomicAdd(&ctr[key], 1) for every active lane, one atomic per distinct key
// of the wave; returns this lane's old value.
__device__ uint32_t wave_atomic_inc(uint32_t* ctr, uint32_t key)
{
	const uint32_t lane = __lane_id();
	uint64_t active = __ballot(1);
	uint32_t result = 0;
	while (active)
	{
		int leader = __ffsll((unsigned long long)active) - 1;
		uint32_t lk = __shfl(key, leader);
		uint64_t mask = __ballot(key == lk) & active;
		uint32_t base = 0;
		if ((int)lane == leader)
			base = atomicAdd(&ctr[lk], (uint32_t)__popcll(mask));
		base = __shfl(base, leader);
		if (key == lk)
			result = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
		active &= ~mask;
	}
	return result;
}

// Record positions of the step: within each slot, entries grouped by point
// count and curvature, from the counts at the step's start (a manifold
// regenerated later in the step keeps its group: any order within a slot is
// valid, the groups only keep waves on one path). Place == false counts the
// groups (bucketCount); Place == true takes positions from the group cursors
// (bucketCursor, set to the group starts in between) and writes the record's
// fixed part: contact, bodies, box-box and curved bits.
//
// Entries are sorted by slot, so a workgroup's entries fall in few slots: the
// lanes of the first kGroupWin slots rank themselves in an LDS table and the
// workgroup adds one global atomic per group it holds (atomics on one address
// from every wave of a large grid would serialize); lanes beyond the window
// fall back to one atomic per wave and group.
constexpr int kGroupThreads = 1024;
constexpr int kGroupWin = 4;

template <bool Place>
__global__ __launch_bounds__(kGroupThreads) void k_giant_group(GiantSolveCtx S)
{
	__shared__ uint32_t tab[kGroupWin * kBuckets];
	__shared__ uint32_t base[kGroupWin * kBuckets];
	__shared__ uint32_t slot0;
	const PassCtx& P = S.X.P;
	const int32_t nValid = (int32_t)S.slotStart[S.counters[CT_SLOTS]];
	const int32_t first = (int32_t)(blockIdx.x * blockDim.x);
	const int32_t e = first + (int32_t)threadIdx.x;
	uint32_t* ctr = Place ? S.bucketCursor : S.bucketCount;
	if (threadIdx.x < (uint32_t)(kGroupWin * kBuckets))
		tab[threadIdx.x] = 0u;
	if (threadIdx.x == 0)
		slot0 = first < nValid ? (uint32_t)(S.keys[first] >> S.bitsI) : 0u;
	__syncthreads();
	const bool valid = e < nValid;
	int32_t c = 0;
	uint32_t key = 0, local = 0;
	int32_t cell = -1;
	if (valid)
	{
		c = (int32_t)S.entries[e];
		key = giant_bucket(P, S, e, c);
		uint32_t off = key / kBuckets - slot0;
		if (off < (uint32_t)kGroupWin)
		{
			cell = (int32_t)(off * kBuckets + key % kBuckets);
			local = atomicAdd(&tab[cell], 1u);
		}
	}
	__syncthreads();
	if (threadIdx.x < (uint32_t)(kGroupWin * kBuckets))
	{
		uint32_t n = tab[threadIdx.x];
		uint32_t b = 0u;
		if (n != 0u)
			b = atomicAdd(&ctr[(slot0 + threadIdx.x / kBuckets) * kBuckets + threadIdx.x % kBuckets], n);
		base[threadIdx.x] = b;
	}
	__syncthreads();
	if (!valid)
		return;
	uint32_t at = cell >= 0 ? base[cell] + local : 0u;
	if (cell < 0)
		at = wave_atomic_inc(ctr, key);
	if (Place)
	{
		S.pos[e] = at;
		S.entryAt[at] = (uint32_t)e;
		S.atOf[c] = at;
		int4 rc = make_rec(P, c, 0);
		rc.w = (rc.w & kRecBoxBox) | ((f2u(P.C.mat[c].w) & CF_CURVED) ? kRecCurved : 0);
		S.rec[at] = rc;
	}
}

// Start of every (slot, group) run; clears the counts for the next step.
__global__ void k_giant_bucket_starts(GiantSolveCtx S)
{
	int32_t s = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (s >= (int32_t)S.counters[CT_SLOTS])
		return;
	uint32_t at = S.slotStart[s];
	for (int b = 0; b < kBuckets; ++b)
	{
		uint32_t n = S.bucketCount[s * kBuckets + b];
		S.bucketCount[s * kBuckets + b] = 0u;
		S.bucketCursor[s * kBuckets + b] = at;
		at += n;
	}
}

// Schedule-ordered copy of a contact's point rows at its record position,
// taken from prepare_contact's registers as it stores them; the rows it does
// not compute (slip, friction anchors) are loaded.
struct CopySink
{
	float4* nrm;
	const DContacts* C;
	int64_t st;
	uint32_t at;
	bool flat, boxBox;
	uint32_t* cnt; // the record's point count, at the entry's sorted position
	mutable float4 slip[4], fa[4], fb[4];
	__device__ void preload(int j, int32_t p) const
	{
		slip[j] = C->pSlip[p];
		if (boxBox)
		{
			fa[j] = C->pFA[p];
			fb[j] = C->pFB[p];
		}
	}
	__device__ void point(int j, int32_t, const float4& a4, const float4& b4, const float4& ja, const float4& jb,
						  const float4& ra, const float4& rb) const
	{
		nrm[(1 + j) * st + at] = a4;
		nrm[(5 + j) * st + at] = b4;
		nrm[(9 + j) * st + at] = ja;
		nrm[(13 + j) * st + at] = jb;
		nrm[(18 + j) * st + at] = slip[j];
		if (flat || (boxBox && j == 0))
		{
			nrm[(22 + j) * st + at] = ra;
			nrm[(26 + j) * st + at] = rb;
		}
		if (boxBox)
		{
			nrm[(30 + j) * st + at] = fa[j];
			nrm[(34 + j) * st + at] = fb[j];
		}
	}
	__device__ void contact(const float4& nc, const float4& mat) const
	{
		*cnt = (uint32_t)f2i(nc.w);
		nrm[at] = nc;
		nrm[17 * st + at] = mat;
	}
};

// Every substep: the contact's constraint rows for the substep, written to
// the contact and to the schedule-ordered copy the three passes read. Lanes
// take the grid contacts in island-list order (ascending contact within an
// island): the contact's columns are read and written coalesced and the
// bodies of neighbouring lanes are near each other; only the copy is
// written at scattered record positions (in record order, lanes of one
// wave would each touch another line of every contact column, since a
// slot holds one contact in every few of an island).
// Workgroup size of prepare: one wave up to kPrepareWaveEntries grid
// contacts, 256 lanes above. A step of a streaming world holds a few tens of
// thousands of grid contacts: 256-lane workgroups would leave most CUs idle
// and put four of the few waves on each busy CU, while single waves spread
// over every CU (the kernel then takes the time of its slowest wave, whose
// manifold regenerations dominate; 0.30 -> 0.27 ms per step on the 200k
// stream). With millions of contacts (a batch of worlds) the larger
// workgroups are faster (6.35 -> 6.08 ms per step).
constexpr int32_t kPrepareWaveEntries = 65536;
__global__ void __launch_bounds__(256) k_giant_prepare(GiantSolveCtx S)
{
	const PassCtx& P = S.X.P;
	const IslandCtx& X = S.X;
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= X.NC || (X.mode[(int32_t)X.entryIsland[k]] & IM_GRID) == 0)
		return;
	int32_t c = (int32_t)X.contactList[k];
	int32_t at = (int32_t)S.atOf[c];
	int4 r = S.rec[at];
	CopySink sink = { S.nrm, &P.C, S.nrmStride, (uint32_t)at, (r.w & kRecCurved) == 0, (r.w & kRecBoxBox) != 0,
					  S.cnt + S.entryAt[at] };
	prepare_contact(P.B, P.C, c, P.h, sink);
}

// Every substep: the record's form (the packet's largest point count, or the
// serial form) and the contact's point count.
// Also counts the substep's entries and points, for the byte model of the
// passes. Runs on reduce_grid workgroups.
__global__ void k_giant_rec_forms(GiantSolveCtx S)
{
	const int32_t nValid = (int32_t)S.slotStart[S.counters[CT_SLOTS]];
	uint32_t entries = 0, pointSum = 0, bb = 0, bbPoints = 0;
	for (int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); e < nValid; e += (int32_t)(gridDim.x * blockDim.x))
	{
		int2 pk = S.packet[e];
		int32_t mp = kRecSerial;
		if (pk.x >= 0)
		{
			mp = 0;
			for (int32_t q = pk.x; q < pk.y; ++q)
			{
				int32_t pc = (int32_t)S.cnt[q];
				mp = pc > mp ? pc : mp;
			}
		}
		uint32_t at = S.pos[e];
		int32_t points = (int32_t)S.cnt[e];
		int32_t w = S.rec[at].w & (kRecBoxBox | kRecCurved);
		S.rec[at].w = w | mp | ((points < 4 ? points : 4) << kRecCountShift);
		entries += 1u;
		pointSum += (uint32_t)points;
		if (w & kRecBoxBox)
		{
			bb += 1u;
			bbPoints += (uint32_t)points;
		}
	}
	block_add<4>(S.counters2, { CT_GIANT_ENTRIES, CT_GIANT_POINTS, CT_GIANT_BB, CT_GIANT_BB_POINTS },
				 { entries, pointSum, bb, bbPoints });
}

// lambdaN of the normal pass back from the schedule-ordered copy to the
// contacts (the stiction and velocity passes and the next steps read it there)
__global__ void k_giant_lambda_back(GiantSolveCtx S)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= (int32_t)S.slotStart[S.counters[CT_SLOTS]])
		return;
	const DContacts& C = S.X.P.C;
	int32_t c = S.rec[k].x;
	const int64_t st = S.nrmStride;
	int32_t count = f2i(S.nrm[k].w);
	for (int j = 0; j < 4 && j < count; ++j)
		C.pJA[pidx(C, c, j)].w = S.nrm[(9 + j) * st + k].w;
}

// `passes` sweeps of one pass kind. The barrier counter is zeroed again by
// the last workgroup to leave, for the next launch.
template <class Pass>
__global__ void __launch_bounds__(kGiantThreads) k_giant_pass(GiantSolveCtx S, int passes)
{
	GridSync grid = GridSync::make(S.barrier, S.X.xcdBar, S.counters2 + CT_GRID_ABORT, S.X.barTimeout);
	const int32_t nSlots = (int32_t)S.counters[CT_SLOTS];
	// slot starts in LDS when they fit
	constexpr int32_t kSlotLds = 2048;
	__shared__ uint32_t sSlots[kSlotLds];
	const uint32_t* ss = S.slotStart;
	if (nSlots + 1 <= kSlotLds)
	{
		for (int32_t i = (int32_t)threadIdx.x; i <= nSlots; i += (int32_t)blockDim.x)
			sSlots[i] = S.slotStart[i];
		__syncthreads();
		ss = sSlots;
	}
	for (int it = 0; it < passes; ++it)
		giant_sweep<Pass>(S, ss, grid, nSlots);
	grid.leave();
}

// The substep's position stage in one cooperative launch: `passes` normal
// sweeps, their lambdas back to the contacts, the stiction sweep and the
// velocity reconstruction of the grid bodies, each stage behind a grid
// barrier (the launch boundaries it replaces cost more than a barrier each).
// The same work as k_giant_pass<NormalsPass>, k_giant_lambda_back,
// k_giant_pass<FrictionPass> and k_giant_bodies(op 1) launched in turn.
__global__ void __launch_bounds__(kGiantThreads) k_giant_position(GiantSolveCtx S, int passes)
{
	GridSync grid = GridSync::make(S.barrier, S.X.xcdBar, S.counters2 + CT_GRID_ABORT, S.X.barTimeout);
	const int32_t nSlots = (int32_t)S.counters[CT_SLOTS];
	constexpr int32_t kSlotLds = 2048;
	__shared__ uint32_t sSlots[kSlotLds];
	const uint32_t* ss = S.slotStart;
	if (nSlots + 1 <= kSlotLds)
	{
		for (int32_t i = (int32_t)threadIdx.x; i <= nSlots; i += (int32_t)blockDim.x)
			sSlots[i] = S.slotStart[i];
		__syncthreads();
		ss = sSlots;
	}
	for (int it = 0; it < passes; ++it)
		giant_sweep<NormalsPass>(S, ss, grid, nSlots);
	const int32_t lanes = (int32_t)(gridDim.x * blockDim.x);
	const int32_t tid = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (S.lambdaBack)
	{
		const DContacts& C = S.X.P.C;
		const int64_t st = S.nrmStride;
		const int32_t nValid = (int32_t)ss[nSlots];
		for (int32_t k = tid; k < nValid; k += lanes)
		{
			int32_t c = S.rec[k].x;
			int32_t count = f2i(S.nrm[k].w);
			for (int j = 0; j < 4 && j < count; ++j)
				C.pJA[pidx(C, c, j)].w = S.nrm[(9 + j) * st + k].w;
		}
		grid.sync();
	}
	giant_sweep<FrictionPass>(S, ss, grid, nSlots);
	{
		const IslandCtx& X = S.X;
		for (int32_t k = tid; k < X.NB; k += lanes)
		{
			int32_t b = (int32_t)X.bodyList[k];
			if (giant_body(X, b))
				reconstruct_velocity(X.P.B, b, X.invH);
		}
	}
	grid.leave();
}

// Co-resident grid size of a cooperative kernel (blocks of kGiantThreads),
// cached per device and kernel: the kernel is a template argument, so each
// kernel has its own cache (the pass kernels share one function type).
template <auto kernel>
static int coop_blocks(World& w)
{
	constexpr int kMaxDevices = 64;
	static std::atomic<int> cache[kMaxDevices];
	bool cacheable = w.device >= 0 && w.device < kMaxDevices;
	if (cacheable)
	{
		int v = cache[w.device].load(std::memory_order_relaxed);
		if (v > 0)
			return v;
	}
	int cus = 0, perCU = 0;
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w.device) != hipSuccess || cus < 1)
		cus = 1;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, kGiantThreads, 0) != hipSuccess || perCU < 1)
		perCU = 1;
	int v = perCU * cus;
	if (cacheable)
		cache[w.device].store(v, std::memory_order_relaxed);
	return v;
}

// Workgroup placement probe: the XCC_ID register of each workgroup.
__global__ void k_xcc_probe(uint32_t* out, int32_t n)
{
	if (threadIdx.x == 0 && (int32_t)blockIdx.x < n)
		out[blockIdx.x] = xcc_id();
}

// Whether grid barriers may take XCD mode on this device (GridSync): over a
// grid of 4096 workgroups the XCC_ID register reads 0 to 7 and every value
// occurs (8 XCDs, each its own L2). Checked once per device;
// M3D_XCD_BARRIER=0 turns the mode off for the process.
constexpr int kXcdMaxDevices = 64;
static std::atomic<int> gXcdState[kXcdMaxDevices]; // 0 unknown, 1 on, 2 off

static bool xcd_barrier_ok(World& w)
{
	if (w.device < 0 || w.device >= kXcdMaxDevices)
		return false;
	int st = gXcdState[w.device].load(std::memory_order_relaxed);
	if (st != 0)
		return st == 1;
	const char* e = getenv("M3D_XCD_BARRIER");
	bool ok = !(e != nullptr && atoi(e) == 0);
	if (ok)
	{
		constexpr int32_t n = 4096;
		uint32_t* d = nullptr;
		std::vector<uint32_t> h(n, 0xFFFFFFFFu);
		ok = hipMalloc(&d, n * sizeof(uint32_t)) == hipSuccess;
		if (ok)
		{
			k_xcc_probe<<<n, 64, 0, w.stream>>>(d, n);
			ok = hipMemcpyAsync(h.data(), d, n * sizeof(uint32_t), hipMemcpyDeviceToHost, w.stream) == hipSuccess &&
				 hipStreamSynchronize(w.stream) == hipSuccess;
			(void)hipFree(d);
		}
		uint32_t seen = 0;
		for (int32_t b = 0; ok && b < n; ++b)
		{
			ok = h[b] < kBarGroups;
			seen |= ok ? 1u << h[b] : 0u;
		}
		ok = ok && seen == (1u << kBarGroups) - 1u;
		(void)hipGetLastError();
	}
	gXcdState[w.device].store(ok ? 1 : 2, std::memory_order_relaxed);
	return ok;
}

// Workgroups of a giant-island launch: enough lanes for every entry, at
// most a cap (a grid barrier costs about 0.04 us more with every workgroup
// that arrives, and a slot's lanes solve one contact each in parallel, so
// lanes beyond the larger slots only add barrier time) and never more than
// can be co-resident. The passes' count comes from the world's tuner.

// The coloring fixpoint does several loads per entry and round, so it
// takes one lane per entry up to kOrderMaxBlocks workgroups (a round is
// bound by its loads: on a 6 M-entry batch 128 workgroups took 2.95 ms per
// step, 512 took 1.39 and 1024 1.35; on a 200 k-entry world 512 and 128
// were even and 1024 slower).
constexpr int kOrderMaxBlocks = 512;

static int env_blocks(const char* name, int dflt)
{
	const char* e = getenv(name);
	int v = e ? atoi(e) : 0;
	return v > 0 ? v : dflt;
}

// Launch of a grid-barrier kernel. Its size never exceeds what every CU's
// occupancy holds at once (giant_blocks, the pass tuner's limit), so the
// workgroups are co-resident with a plain launch, whose dispatch follows the
// stream's previous kernel directly; a cooperative launch (which also
// guarantees it) first waits about 11 us for the runtime's switch of queues
// on each side. A world whose barrier wait ever exceeded its bound
// (CT_GRID_ABORT) uses cooperative launches from then on, as does every
// world with M3D_COOP_LAUNCH=1.
static hipError_t grid_launch(World& w, const void* kernel, int blocks, void** args, hipStream_t s)
{
	static const bool forceCoop = env_blocks("M3D_COOP_LAUNCH", 0) == 1;
	if (w.coopLaunch || forceCoop)
		return hipLaunchCooperativeKernel(kernel, dim3(blocks), dim3(kGiantThreads), args, 0, s);
	return hipLaunchKernel(kernel, dim3(blocks), dim3(kGiantThreads), args, 0, s);
}

template <auto kernel>
static int giant_blocks(World& w, int32_t entries, int cap)
{
	int b = (int)std::min<int64_t>(((int64_t)entries + kGiantThreads - 1) / kGiantThreads, cap);
	return std::max(1, std::min(b, coop_blocks<kernel>(w)));
}

// Captures `body`'s launches on the world's stream into a graph and launches
// it: the world's executable graph is updated in place when the topology
// matches (kernel arguments and grid sizes change from step to step) and
// instantiated again otherwise. A capture the runtime refuses falls back to
// direct launches (and turns graphs off for the world).
template <class Body>
static bool run_step_graph(World& w, hipStream_t s, Body&& body)
{
	if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
	{
		(void)hipGetLastError();
		w.stepGraphs = false;
		return body();
	}
	bool ok = body();
	hipGraph_t g = nullptr;
	hipError_t e = hipStreamEndCapture(s, &g);
	if (!ok || e != hipSuccess || g == nullptr)
	{
		if (g != nullptr)
			(void)hipGraphDestroy(g);
		(void)hipGetLastError();
		w.stepGraphs = false;
		return body();
	}
	if (w.stepGraphExec != nullptr)
	{
		hipGraphNode_t errNode = nullptr;
		hipGraphExecUpdateResult res;
		if (hipGraphExecUpdate(w.stepGraphExec, g, &errNode, &res) != hipSuccess)
		{
			(void)hipGetLastError();
			(void)hipGraphExecDestroy(w.stepGraphExec);
			w.stepGraphExec = nullptr;
		}
	}
	if (w.stepGraphExec == nullptr && hipGraphInstantiate(&w.stepGraphExec, g, nullptr, nullptr, 0) != hipSuccess)
	{
		(void)hipGetLastError();
		w.stepGraphExec = nullptr;
		(void)hipGraphDestroy(g);
		w.stepGraphs = false;
		return body();
	}
	(void)hipGraphDestroy(g);
	++w.stepGraphLaunches;
	return hip_ok(hipGraphLaunch(w.stepGraphExec, s), "step graph");
}

bool solve_islands(World& w, const IslandLists& L, float dt, int substeps)
{
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	IslandCtx X;
	X.P.B = make_body_view(w);
	X.P.C = make_contact_view(w);
	X.P.h = dt / (float)substeps;
	X.P.gravityLen = m3d_length(w.def.gravity);
	X.P.maxCorrection = kMaxDepenetrationSpeed * X.P.h;
	X.bodyList = L.bodyList;
	X.bodyStart = L.bodyStart;
	X.contactList = L.contactList;
	X.contactStart = L.contactStart;
	X.entryIsland = L.entryIsland;
	X.mode = L.mode;
	X.islandOf = L.islandOf;
	X.J = w.djoints.p;
	X.nJ = w.djointCount;
	X.jointIsland = L.jointIsland;
	X.contactColor = w.contactColor.p;
	X.gravity = w.def.gravity;
	X.invH = X.P.h > 0.0f ? 1.0f / X.P.h : 0.0f;
	X.iterations = w.def.positionIterations > 0 ? w.def.positionIterations : 2;
	X.substeps = substeps;
	X.I = L.I;
	X.NB = L.NB;
	X.NC = L.NC;
	X.xcdBar = w.xcdBarrier && xcd_barrier_ok(w) ? 1u : 0u;
	X.barTimeout = w.barTimeout;

	if (L.NB == 0)
		return true;
	// joints: levels of their serial order (joint_levels, step.cpp)
	const bool joints = w.djointCount > 0 && w.jointLevelStart.size() > 1;
	auto joint_levels = [&](int op) {
		w.mark_begin(M3DX_K_JOINT_ISLANDS);
		for (size_t l = 0; l + 1 < w.jointLevelStart.size(); ++l)
		{
			int32_t b0 = w.jointLevelStart[l], b1 = w.jointLevelStart[l + 1];
			k_joint_level<<<gridk(b1 - b0), 256, 0, s>>>(X, w.jointOrder.p, b0, b1, op);
		}
		w.mark_end();
	};
	auto joint_reset = [&]() {
		if (joints)
			k_joints_all<<<gridk(w.djointCount), 256, 0, s>>>(X, 3);
	};
	if (joints)
	{
		w.mark_begin(M3DX_K_JOINT_ISLANDS);
		k_joint_island_inertia<<<gridk(L.NB), 256, 0, s>>>(X);
		k_joints_all<<<gridk(w.djointCount), 256, 0, s>>>(X, 0);
		w.mark_end();
	}

	GiantSolveCtx S;
	S.lambdaBack = 1;
	S.X = X;
	if (L.NC == 0)
	{
		// no contacts: bodies and joints only
		for (int sub = 0; sub < substeps; ++sub)
		{
			w.mark_begin(M3DX_K_GIANT_BODIES);
			k_giant_bodies<<<gridk(L.NB), 256, 0, s>>>(S, 0);
			w.mark_end();
			if (joints)
			{
				joint_reset();
				for (int it = 0; it < X.iterations; ++it)
					joint_levels(1);
			}
			w.mark_begin(M3DX_K_GIANT_BODIES);
			k_giant_bodies<<<gridk(L.NB), 256, 0, s>>>(S, 1);
			w.mark_end();
			if (joints)
				joint_levels(2);
		}
		return true;
	}

	// ---- the grid path ----
	int32_t NC = L.NC;
	size_t two = 2 * (size_t)NC;
	int bitsN = 1;
	while ((1ll << bitsN) <= (int64_t)n)
		++bitsN;
	int32_t maxS = NC + kMaxGraphColors + 1; // slot bound: 8 colors + overflow levels
	int bitsS = 1;
	while ((1ll << bitsS) <= (int64_t)maxS)
		++bitsS;
	if (!w.s64a.ensure(two) || !w.s64b.ensure(two) || !w.s32a.ensure(two) || !w.s32b.ensure(two) ||
		!w.s32c.ensure((size_t)n + 2) || !w.sColor.ensure(NC) || !w.sLevel.ensure(NC) ||
		!w.sSlotStart.ensure((size_t)maxS + 1) || !w.sPacket.ensure(NC) || !w.sRecG.ensure(NC) || !w.sPosG.ensure(3 * (size_t)NC) || !w.sAtOf.ensure((size_t)std::max(w.contactCount, 1)) ||
		!w.sNrm.ensure(kSchedBlocks * (size_t)NC) || !w.sPred.ensure(NC) || !w.sColored.ensure(NC) ||
		!w.sBucket.ensure(2 * (size_t)(maxS + 1) * kBuckets) || !w.sCtr.ensure(16) || !w.gridBar.ensure(2 * (size_t)kBarWords) || !w.sDirty.ensure(two + 4) ||
		!ensure_prim_scratch(w, std::max<int64_t>((int64_t)two + n + 2, NC)))
		return false;
	uint32_t* bodyCount = w.s32c.p;
	launch_fill_u32(bodyCount, 0u, (int64_t)n + 1, s);
	k_giant_adj<<<gridk(NC), 256, 0, s>>>(X, w.sPred.p, w.sColored.p, w.sColor.p, w.sLevel.p, n, w.s64a.p, w.s32a.p,
										  bodyCount);
	scan_exclusive(bodyCount, bodyCount, (int64_t)n + 1, nullptr, w.scanStatus.p, s);
	// The two sorts below leave an odd pass count's result in their
	// ping-pong buffers (no copy back): the adjacency lands in one pair of
	// the (s64, s32) buffers and the slot keys go to the other, which the
	// slot sort then ping-pongs with the adjacency's pair (free once the
	// fixpoints are done).
	const bool adjAlt = radix_sort_pairs_in_place_or_alt(w.s64a.p, w.s32a.p, w.s64b.p, w.s32b.p, (int64_t)two, bitsN,
														 w.primScratch.p, w.scanStatus.p, s);
	uint64_t* adjK = adjAlt ? w.s64b.p : w.s64a.p;
	uint32_t* adjV = adjAlt ? w.s32b.p : w.s32a.p;
	uint64_t* slotK = adjAlt ? w.s64a.p : w.s64b.p;
	uint32_t* slotV = adjAlt ? w.s32a.p : w.s32b.p;
	k_giant_pred<<<gridk((int64_t)two), 256, 0, s>>>(X, adjK, adjV, bodyCount, (int64_t)two, n, w.sPred.p);
	{
		FillSegs F;
		F.add(w.sCtr.p, 0u, 16);                                   // fixpoint counters [0, 6)
		F.add(w.gridBar.p, 0u, 2 * (int64_t)kBarWords);            // the passes' and the order's barriers
		F.add((uint32_t*)w.sDirty.p, 0u, (int64_t)(two + 3) / 4); // fixpoint worklists
		launch_fill_multi(F, s);
	}

	GiantCtx G;
	G.X = X;
	G.adjStart = bodyCount;
	G.adj = adjV;
	G.pred = w.sPred.p;
	G.colored = w.sColored.p;
	G.color = w.sColor.p;
	G.level = w.sLevel.p;
	G.ctr = w.sCtr.p;
	G.bar = w.gridBar.p + kBarWords;
	G.counters = w.counters.p;
	G.slotKeys = slotK;
	G.slotVals = slotV;
	G.dirty = w.sDirty.p;
	G.bitsI = L.bitsI;
	G.maxS = maxS;
	if (w.subCount > 0)
	{
		// batches: one workgroup per sub-world
		if (!w.subEntryStart.ensure((size_t)w.subCount + 1))
			return false;
		k_sub_entry_start<<<gridk((int64_t)NC + 1), 256, 0, s>>>(G, w.bodySub.p, w.subCount, w.subEntryStart.p);
		w.mark_begin(M3DX_K_GIANT_ORDER);
		k_giant_order_sub<<<(unsigned)w.subCount, kGiantThreads, 0, s>>>(G, w.subEntryStart.p);
		w.mark_end();
	}
	else
	{
		static const int cap = env_blocks("M3D_ORDER_BLOCKS", kOrderMaxBlocks);
		int blocks = giant_blocks<k_giant_order>(w, NC, cap);
		void* args[] = { &G };
		w.mark_begin(M3DX_K_GIANT_ORDER);
		hipError_t err = grid_launch(w, (const void*)k_giant_order, blocks, args, s);
		w.mark_end();
		if (!hip_ok(err,
					"giant order"))
			return false;
	}
	if (radix_sort_pairs_in_place_or_alt(slotK, slotV, adjK, adjV, NC, L.bitsI + bitsS + 1, w.primScratch.p,
										 w.scanStatus.p, s))
	{
		slotK = adjK;
		slotV = adjV;
	}
	k_giant_slot_starts<<<gridk(maxS + 1), 256, 0, s>>>(slotK, NC, L.bitsI, maxS, w.sSlotStart.p);
	k_giant_packets<<<gridk(NC), 256, 0, s>>>(slotK, NC, L.bitsI, w.sPacket.p);

	S.entries = slotV;
	S.packet = w.sPacket.p;
	S.slotStart = w.sSlotStart.p;
	S.counters = w.counters.p;
	S.counters2 = w.counters.p;
	S.barrier = w.gridBar.p;
	S.rec = w.sRecG.p;
	S.nrm = w.sNrm.p;
	S.nrmStride = NC;
	S.keys = slotK;
	S.bitsI = L.bitsI;
	S.bucketCount = w.sBucket.p;
	S.bucketCursor = w.sBucket.p + (size_t)(maxS + 1) * kBuckets;
	S.pos = w.sPosG.p;
	S.cnt = w.sPosG.p + NC;
	S.entryAt = w.sPosG.p + 2 * (size_t)NC;
	S.atOf = w.sAtOf.p;
	launch_fill_u32(S.bucketCount, 0u, (int64_t)(maxS + 1) * kBuckets, s);
	const unsigned groupGrid = (unsigned)((NC + kGroupThreads - 1) / kGroupThreads);
	k_giant_group<false><<<groupGrid, kGroupThreads, 0, s>>>(S);
	k_giant_bucket_starts<<<gridk(maxS), 256, 0, s>>>(S);
	k_giant_group<true><<<groupGrid, kGroupThreads, 0, s>>>(S);
	// more entries per slot pay for the dearer barrier of a wider grid: about
	// one workgroup per 1536 entries, 64 to 256 (measured on a 190 k-entry
	// stream step and a 4 M-entry batch step)
	// workgroup count from the world's pass tuner (M3D_GIANT_BLOCKS or
	// m3dx_set_pass_blocks force one)
	static const int capEnv = env_blocks("M3D_GIANT_BLOCKS", 0);
	if (capEnv > 0 && w.tuner.forcedBlocks == 0)
		w.tuner.forcedBlocks = capEnv;
	const int coopMax = std::min(std::min(coop_blocks<k_giant_pass<NormalsPass>>(w), coop_blocks<k_giant_position>(w)),
								 std::min(coop_blocks<k_giant_pass<FrictionPass>>(w),
										  coop_blocks<k_giant_pass<VelocityPass>>(w)));
	bool measure = false;
	int blocks = w.tuner_pick(NC, coopMax, measure);
	w.tuner.lastBlocks = blocks;
	if (measure)
		(void)hipEventRecord(w.tuner.ev0, s);
	// a count the runtime refuses for a cooperative launch (its limit can be
	// below the occupancy estimate) is lowered until accepted, and the world
	// keeps the accepted count as its limit
	auto coop = [&](const void* kernel, int passes) {
		void* args[] = { &S, &passes };
		for (;;)
		{
			hipError_t e = grid_launch(w, kernel, blocks, args, s);
			if (e != hipErrorCooperativeLaunchTooLarge || blocks <= 1)
				return hip_ok(e, "giant pass");
			(void)hipGetLastError();
			blocks = std::max(1, blocks * 3 / 4);
			w.tuner.coopLimit = blocks;
			w.tuner.lastBlocks = blocks;
			w.tuner.pendingFailed = measure;
		}
	};
	const int prepThreads = NC <= kPrepareWaveEntries ? 64 : 256;
	auto substepLoop = [&]() -> bool {
	for (int sub = 0; sub < substeps; ++sub)
	{
		S.lambdaBack = sub == substeps - 1 ? 1 : 0;
		w.mark_begin(M3DX_K_GIANT_BODIES);
		k_giant_bodies<<<gridk(L.NB), 256, 0, s>>>(S, 0);
		w.mark_end();
		w.mark_begin(M3DX_K_GIANT_PREPARE);
		k_giant_prepare<<<(unsigned)((NC + prepThreads - 1) / prepThreads), prepThreads, 0, s>>>(S);
		w.mark_end();
		k_giant_rec_forms<<<reduce_grid(NC, 256), 256, 0, s>>>(S);
		if (!diff_record_substep(w, sub, S.atOf, S.rec))
			return false;
		bool ok = true;
		// one launch for the position stage unless joints interleave with the
		// normal sweeps or the passes are timed one by one
		if (!joints && !w.timing)
		{
			ok = coop((const void*)k_giant_position, X.iterations);
			if (!ok)
				return false;
		}
		else
		{
		if (joints)
		{
			// each position iteration: the joints, then the contacts
			joint_reset();
			for (int it = 0; it < X.iterations && ok; ++it)
			{
				joint_levels(1);
				w.mark_begin(M3DX_K_GIANT_NORMALS);
				ok = coop((const void*)k_giant_pass<NormalsPass>, 1);
				w.mark_end();
			}
		}
		else
		{
			w.mark_begin(M3DX_K_GIANT_NORMALS);
			ok = coop((const void*)k_giant_pass<NormalsPass>, X.iterations);
			w.mark_end();
		}
		if (S.lambdaBack)
			k_giant_lambda_back<<<gridk(NC), 256, 0, s>>>(S);
		w.mark_begin(M3DX_K_GIANT_FRICTION);
		ok = ok && coop((const void*)k_giant_pass<FrictionPass>, 1);
		w.mark_end();
		if (!ok)
			return false;
		w.mark_begin(M3DX_K_GIANT_BODIES);
		k_giant_bodies<<<gridk(L.NB), 256, 0, s>>>(S, 1);
		w.mark_end();
		}
		w.mark_begin(M3DX_K_GIANT_VELOCITY);
		ok = coop((const void*)k_giant_pass<VelocityPass>, 1);
		w.mark_end();
		if (!ok)
			return false;
		if (joints)
			joint_levels(2);
	}
	return true;
	};
	// A batch runs the loop as one graph once its workgroup count is settled
	// and was accepted by a direct cooperative launch; never while the tuner
	// measures, kernels are timed or steps are recorded.
	bool graph = w.subCount > 0 && w.stepGraphs && !measure && !w.timing && blocks == w.validatedBlocks &&
				 !diff_is_recording(w);
	bool ok = graph ? run_step_graph(w, s, substepLoop) : substepLoop();
	if (ok && !graph)
		w.validatedBlocks = blocks;
	if (measure)
		(void)hipEventRecord(w.tuner.ev1, s);
	return ok;
}

} // namespace mx
