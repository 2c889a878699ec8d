// Host-side world of the device backend: the host mirror of bodies and
// joints that serves the m3d_* API, the device-resident SoA state the step
// kernels run on, and the bookkeeping that keeps the two consistent.
//
// Ownership: the device copy of body state is authoritative after a step;
// the host mirror is refreshed lazily (one bulk download on the first API
// access after a step). API writes edit the host mirror and mark the body
// dirty; dirty bodies are scattered to the device at the start of the next
// step. Contacts live only on the device; destroys between steps are
// resolved on the host against a downloaded contact sidecar and applied to
// the device contact arrays as one permutation at the next step.
#pragma once

#include "mx_solver.h"
#include "mx_prims.h"
#include "melon3d_ext.h"

#include <string>
#include <vector>

namespace mx
{

void set_error(const char* fmt, ...);
bool hip_ok(hipError_t e, const char* what);

// Device buffer with capacity-doubling growth.
template <class T>
struct DBuf
{
	T* p = nullptr;
	size_t cap = 0;
	// Ensures room for n elements; keeps the first `keep` elements.
	bool ensure(size_t n, size_t keep = 0)
	{
		if (n <= cap)
		{
			return true;
		}
		size_t nc = cap ? cap : 64;
		while (nc < n)
		{
			nc = nc * 3 / 2 + 64;
		}
		T* np = nullptr;
		if (hipMalloc((void**)&np, nc * sizeof(T)) != hipSuccess)
		{
			set_error("hipMalloc of %zu bytes failed", nc * sizeof(T));
			return false;
		}
		// new memory starts zeroed (the scan's tagged status words rely on
		// never meeting a stale tag); allocations are rare
		(void)hipMemset(np, 0, nc * sizeof(T));
		(void)hipDeviceSynchronize();
		if (p != nullptr)
		{
			if (keep > 0)
			{
				(void)hipMemcpy(np, p, keep * sizeof(T), hipMemcpyDeviceToDevice);
			}
			(void)hipFree(p);
		}
		p = np;
		cap = nc;
		return true;
	}
	void release()
	{
		if (p != nullptr)
		{
			hipFree(p);
		}
		p = nullptr;
		cap = 0;
	}
};

struct HostBody
{
	m3d_vec3 position;
	m3d_quat rotation;
	m3d_vec3 linearVelocity;
	m3d_vec3 angularVelocity;
	m3d_vec3 prevPosition;
	m3d_quat prevRotation;
	m3d_vec3 prevLinearVelocity;
	m3d_vec3 prevAngularVelocity;
	float invMass;
	m3d_vec3 invInertiaLocal;
	m3d_mat3 invInertiaWorld;
	Shape shape;
	float friction;
	float restitution;
	float rollingResistance;
	float mass;
	m3d_vec3 force;
	m3d_vec3 torque;
	m3d_filter filter;
	bool isSensor;
	m3d_body_type type;
	float linearDamping;
	float angularDamping;
	float gravityScale;
	float sleepTime;
	float smoothedEnergy;
	bool enableSleep;
	bool isAwake;
	bool inUse;
	uint32_t generation;
	uint64_t userData;
	m3d_aabb fat;
	uint8_t bp; // BP_HAS_FAT | BP_IN_STABLE_TIER | BP_IN_STABLE_CELLS persist across steps
};

// Per-body column sets used for bulk transfers.
struct BodyColumns
{
	DBuf<float4> pos, rot, vel, angVel, prevPos, prevRot, prevVel, prevAngVel, invIL, iw0, iw1, iw2, shape, mat,
		force, torque, fatLo, fatHi;
	DBuf<uint32_t> flags;
	DBuf<uint4> filter;
	DBuf<uint8_t> bp, sb;
};

struct ContactColumns
{
	DBuf<int2> bodies;
	DBuf<int2> hint; // last greedy color and overflow level (giant islands), a warm start
	DBuf<float4> mat, normal;
	DBuf<float4> cache; // pose cache, 4 per contact
	DBuf<float4> pA, pB, pJA, pJB, pRA, pRB, pFA, pFB, pSlip; // 4 * cap each
};

struct StepCounters; // device-side counters block (k_step.hip)
struct DiffState;    // recorded steps of a differentiable world (k_diff.hip)
void diff_release(struct World& w);

struct World
{
	m3d_world_def def;
	int device = 0;
	hipStream_t stream = nullptr;

	// ---- host mirror ----
	std::vector<HostBody> bodies;
	std::vector<int32_t> freeBodies;
	std::vector<DJoint> joints;
	std::vector<uint32_t> jointGen;
	std::vector<int32_t> freeJoints;
	bool gridDirty = true;
	bool stableDirty = true;
	bool stableForce = false;
	int32_t stepsSinceStableBuild = 1000;
	bool anyForcesApplied = false;
	bool jointedPairsDirty = true;
	std::vector<m3d_contact_event> beginEvents, endEvents;
	m3d_profile profile{};

	// ---- sync state ----
	bool hostStale = false;          // device bodies newer than the host mirror
	bool queryRefreshed = false;     // the grid's last refresh was a query's (dt = 0), after the last step moved bodies
	int32_t gridBodies = 0;          // body slots the grid tiers were last built over
	// scene queries: bodies the grid may not find (see k_query.hip), query
	// results before their sort
	DBuf<int32_t> qExtra;
	DBuf<uint32_t> qFlag, qOffs;
	DBuf<uint32_t> qUFlag, qUOffs; // repeat removal (qFlag, qOffs hold the extra list)
	DBuf<uint64_t> qPairs, qPairsAlt;
	DBuf<uint32_t> qVals, qValsAlt;
	std::vector<int32_t> dirtyList;  // bodies to upload at the next step
	std::vector<uint8_t> dirtyFlag;
	bool jointsDirty = true;         // joint table to upload
	std::vector<uint8_t> materialRefresh; // bodies whose contacts recombine materials
	bool anyMaterialRefresh = false;

	// Contact sidecar mirror for between-step destroys, in pinned memory the
	// device fills at the end of a step: per contact (body a, body b, flags,
	// original slot); per body the current positions of its contacts
	// (hcList[hcStart[b] .. hcStart[b + 1]), -1 once removed); per contact its
	// two entries in hcList (hcSlot).
	bool sidecarValid = false;
	int4* hcSide = nullptr;
	int32_t hcCount = 0;        // contacts in the mirror
	int32_t hcN = 0;            // bodies the per-body index covers
	int32_t* hcStart = nullptr;
	int32_t* hcList = nullptr;
	int2* hcSlot = nullptr;
	std::vector<int32_t> hcMoved; // positions that received another contact
	int32_t* pinnedCsr = nullptr;
	size_t pinnedCsrCap = 0;
	int32_t sidecarPrefetchN = 0;
	DBuf<int32_t> dCsr;
	DBuf<uint64_t> csrKeys, csrKeysAlt; // (body, side) pairs sorted into the contact index
	DBuf<uint32_t> csrVals, csrValsAlt;
	// bodies to wake on the device at the next step (partners of destroyed
	// bodies whose host record is not current)
	// packed sidecar (body a, body b, flags) copied to pinned memory at the
	// end of a step when the caller destroyed bodies in the last window
	bool destroyedSinceStep = false;
	bool sidecarPrefetched = false;
	int32_t sidecarPrefetchCount = 0;
	int4* pinnedSidecar = nullptr;
	size_t pinnedSidecarCap = 0;
	hipEvent_t sidecarReady = nullptr;
	DBuf<int4> dSidecar;
	std::vector<int32_t> deviceWake;
	std::vector<int32_t> wakeUpload; // the list being uploaded (kept until the next push)
	DBuf<int32_t> dWake;
	bool pendingPermutation = false;
	// bulk despawn: per-body destroy rank (INT32_MAX between calls), the
	// call's (slot, rank) list, and the contacts touching those bodies
	DBuf<int32_t> respRank;
	DBuf<int2> respList;
	DBuf<int4> respOut;
	int4* pinnedResp = nullptr;
	size_t pinnedRespCap = 0;
	std::vector<int32_t> respRankHost; // per body slot, -1 between calls
	// pinned staging of the dirty-body upload, reusable once uploadDone fired
	char* pinnedUpload = nullptr;
	size_t pinnedUploadCap = 0;
	hipEvent_t uploadDone = nullptr;
	// the step's touch transitions (body a, body b, kind), waiting for events
	int4* pinnedTrans = nullptr;
	size_t pinnedTransCap = 0;
	hipEvent_t transReady = nullptr;
	int32_t pendingTrans = 0;
	bool transInFlight = false; // listed on the device, count not yet read back
	DBuf<int32_t> transList;    // contacts with a transition, ascending
	DBuf<int4> transInfo;       // (body a, body b, kind, contact) per transition
	// pinned staging of the contact removal (flags, end events, order)
	char* pinnedRemoval = nullptr;
	size_t pinnedRemovalCap = 0;
	int32_t removalSpec = 0; // removed contacts the staging expects (twice the last count)
	// pinned staging of the coarse-list candidate segments
	char* pinnedSegs = nullptr;
	size_t pinnedSegsCap = 0;

	// ---- device state ----
	int32_t deviceBodyCap = 0;  // capacity of the body columns
	int32_t deviceBodyCount = 0; // bodies.size() as known by the device
	BodyColumns bc;
	ContactColumns cc;
	int32_t contactCap = 0;
	int32_t contactCount = 0;
	DBuf<DJoint> djoints;
	int32_t djointCount = 0;
	// joints in levels of their serial order (index order): level l is
	// jointOrder[jointLevelStart[l] .. jointLevelStart[l + 1])
	std::vector<int32_t> jointLevelStart;
	DBuf<int32_t> jointOrder;
	DBuf<uint64_t> jointedKeys;
	int32_t jointedCount = 0;

	// grid tiers
	DBuf<uint64_t> activeKeys, stableKeys, sortKeysAlt, stableKeysNext;
	DBuf<uint32_t> activeBodies, stableBodies, sortValsAlt, stableBodiesNext;
	int64_t activeCount = 0, stableCount = 0;
	std::vector<int32_t> activeOversized, stableOversized; // host copies (small)
	DBuf<int32_t> dActiveOversized, dStableOversized;

	// scratch
	DBuf<uint32_t> u32a, u32b, u32c, u32d;
	DBuf<uint64_t> u64a, u64b, u64c;
	DBuf<int32_t> i32a, i32b, i32c, i32d, i32e;
	DBuf<uint8_t> u8a;
	DBuf<char> primScratch;       // radix sort histograms
	DBuf<uint64_t> scanStatus;    // scan look-back words (nothing else writes them)
	DBuf<uint32_t> pairRuns;      // pair enumeration: run tables, item counts and offsets
	DBuf<uint32_t> boundParts;    // tier builds: per-workgroup cell bounds (six words each)
	DBuf<uint64_t> overSegs;      // coarse-list candidate segments
	DBuf<uint64_t> hashKeys;
	DBuf<uint32_t> hashVals;
	DBuf<int32_t> islandOf;       // per body: island id (-1)
	DBuf<uint32_t> islandAwakeD;  // per island: any body awake
	DBuf<uint32_t> islandMode;    // per island: awake / colored / has joints
	// parallel solve scratch
	DBuf<uint64_t> s64a, s64b;
	DBuf<uint32_t> s32a, s32b, s32c, sSlotStart;
	DBuf<int32_t> sColor, sLevel, sJointList;
	DBuf<uint32_t> sCtr;          // convergence counters of the giant-island fixpoints
	DBuf<int2> sPacket;
	DBuf<int4> sRecG;             // grid path: solve records by schedule position
	DBuf<int32_t> tailOrder;      // removal: sources of the holes, in the reference's order
	DBuf<uint32_t> gridBar;       // grid path: barrier counters of the passes and of the ordering
	DBuf<uint32_t> sPosG;         // grid path: record position and point count by sorted entry, entry by position
	DBuf<uint32_t> sAtOf;         // grid path: record position by contact
	DBuf<float4> sNrm;            // grid path: normal-pass inputs by schedule position (17 x entries)
	DBuf<uint32_t> sBucket;       // giant islands: per (slot, group) counts and cursors
	DBuf<int4> sPred;             // giant islands: predecessor ranges per entry
	DBuf<uint8_t> sColored;       // grid path: entry's island is colored
	DBuf<uint8_t> sDirty;         // grid path: fixpoint worklist flags (2 x entries)
	DBuf<int32_t> contactColor;   // per contact: -1 / 0..8
	int32_t lastIslandCount = 0;
	bool lastHadIslands = false;
	DBuf<uint32_t> counters;      // device counters block
	uint32_t* hostCounters = nullptr; // pinned mirror of the counters block
	DBuf<int2> moves;             // permutation moves (dst, src)
	DBuf<int32_t> eventIdx;

	m3dx_step_stats stats{};

	// ---- batched sub-worlds (subCount 0: a plain world) ----
	// Bodies of sub-world s occupy [subBodyStart[s], subBodyStart[s + 1]);
	// contacts stay contiguous by sub-world, each range in the order a world
	// of its own would hold them, so every sub-world steps exactly as it
	// would alone.
	int32_t subCount = 0;
	int32_t subCurrent = 0;
	std::vector<int32_t> subBodyStart;  // subCount + 1
	std::vector<int32_t> bodySubHost;   // per body
	std::vector<std::vector<int32_t>> subFreeBodies; // per sub-world: its destroyed slots (LIFO)
	bool bodySubDirty = false;
	std::vector<uint8_t> subDirtyHost;  // stable-tier dirt raised by API calls
	bool anySubDirtyHost = false;
	DBuf<int32_t> bodySub;              // device copy of bodySubHost
	DBuf<int32_t> subBodyStartD;        // device copy of subBodyStart
	DBuf<int32_t> subEntryStart;        // grid solve of a batch: first solver entry of each sub-world
	int32_t maxSubBodies = 0;           // largest sub-world's body count
	// end events of a batch step's removals: flags, offsets, indices, gathered events
	DBuf<uint32_t> bEvFlag, bEvOff;
	DBuf<int32_t> bEvIdx;
	DBuf<int4> bEvOut;
	DBuf<uint32_t> subStable;           // per sub-world SUB_* bits | steps since build << 8
	DBuf<uint8_t> subDirtyUpload;
	DBuf<int32_t> subContactStart;      // subCount + 1, contacts by sub-world
	DBuf<int32_t> subWork;              // per sub-world counters of contact placement
	DBuf<uint64_t> newKeys;             // created pairs of this step, grouped by sub-world
	int32_t newCount = 0;
	std::vector<uint32_t> subStableHost;
	ContactColumns cc2;                 // placement target, swapped with cc
	int32_t contactCap2 = 0;

	DiffState* diff = nullptr;          // recorded steps for the adjoint

	// ---- kernel timing (off unless enabled) ----
	struct KernelMark
	{
		int kernel;
		hipEvent_t begin, end;
		double bytes;
	};
	bool timing = false;
	std::vector<KernelMark> marks;     // events of the current step
	std::vector<hipEvent_t> eventPool; // reused across steps
	size_t eventsUsed = 0;
	m3dx_kernel_time ktime[M3DX_K_COUNT]{};
	// records the begin event of kernel k (no-op unless timing)
	void mark_begin(int k, double bytes = 0.0);
	void mark_end();
	void collect_marks(); // after the step's final synchronization

	// Workgroup count of the giant-island passes, chosen by measurement: per
	// size class of the step's solver entries (floor log2), every candidate
	// is timed on kTunerRounds steps (pass time per slot, HIP events around
	// the substep loop) and the fastest is kept. The count changes no result
	// (a slot's contacts touch disjoint bodies; the order inside a slot is
	// free). forcedBlocks > 0 (API or M3D_GIANT_BLOCKS) bypasses it.
	struct PassTuner
	{
		static constexpr int kClasses = 40;
		static constexpr int kCands = 6;
		static constexpr int kRounds = 2;
		struct Class
		{
			int trials = 0;            // measurements taken
			int chosen = 0;            // workgroups once tuned (0: tuning)
			float perSlot[kCands] = {}; // best ms per slot of each candidate
		};
		Class cls[kClasses];
		int forcedBlocks = 0;
		int coopLimit = 0; // largest count a cooperative launch accepted after a refusal (0: none seen)
		bool pendingFailed = false;
		int pendingClass = -1, pendingCand = -1;
		hipEvent_t ev0 = nullptr, ev1 = nullptr;
		int lastBlocks = 0;
	} tuner;
	static constexpr int kTunerCands[PassTuner::kCands] = { 32, 64, 128, 192, 256, 384 };
	// picks the passes' workgroup count for a step of `entries` solver entries
	// (coopMax: co-resident limit); a pick being measured records ev0 / ev1
	int tuner_pick(int32_t entries, int coopMax, bool& measure);
	// after the step's final synchronization: the measurement of a pick
	void tuner_step_end(int32_t slots);
	int32_t stepFailures = 0; // steps that stopped on an error (m3dx_last_error)
	// the giant substep loop of a batch as one hipGraph (captured each step,
	// the executable graph updated in place)
	bool stepGraphs = true;
	// grid barriers of the solver passes in XCD mode when the device allows
	// it (one L2 write-back and invalidation per XCD; k_isolve.hip GridSync)
	bool xcdBarrier = true;
	// grid-barrier kernels: cooperative launches (else plain launches of a
	// co-resident size), and the longest barrier wait in wall-clock ticks
	bool coopLaunch = false;
	uint32_t barTimeout = 1u << 28;
	int validatedBlocks = 0; // workgroups a direct cooperative launch accepted
	hipGraphExec_t stepGraphExec = nullptr;
	int64_t stepGraphLaunches = 0;

	World();
	~World();
};

// ---- host <-> device transfer (world_sync.cpp) ----
bool ensure_device_bodies(World& w, int32_t n);
bool ensure_device_contacts(World& w, int32_t n);
// grows one contact column set to at least n contacts, keeping the first keep
bool ensure_contact_columns(World& w, ContactColumns& c, int32_t& cap, int32_t n, size_t keep);
DContacts columns_view(ContactColumns& cc, int32_t cap);
bool upload_dirty(World& w);
bool flush_transitions(World& w);
bool download_bodies(World& w);
void ensure_host(World& w);
void mark_dirty(World& w, int32_t i);
bool ensure_sidecar(World& w);
bool apply_pending_permutation(World& w);

// ---- step (k_step.hip) ----
void world_step(World& w, float dt, int substeps);
bool refresh_broadphase(World& w, float dt);
void wake_body_host(World& w, int32_t i);
void set_stable_dirty(World& w, int32_t i);

} // namespace mx

struct m3d_world
{
	mx::World w;
};
