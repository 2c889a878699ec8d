// Launch wrappers of the step stages (one per pipeline stage), called by the
// host orchestrator in step.cpp.
#pragma once

#include "mx_world.h"

namespace mx
{

// device counters block layout
enum Counter : int
{
	CT_ENTRIES = 0,
	CT_OVERSIZED,
	CT_UX_MIN,
	CT_UX_MAX,
	CT_UY_MIN,
	CT_UY_MAX,
	CT_UZ_MIN,
	CT_UZ_MAX,
	CT_CANDIDATES,
	CT_NEW_CONTACTS,
	CT_REMOVED,
	CT_HOLES,
	CT_TAILS,
	CT_END_EVENTS,
	CT_TRANSITIONS,
	CT_REGEN,
	CT_CHANGED,
	CT_ANY_AWAKE_DYN,
	CT_ISLANDS,
	CT_TOUCHING,
	CT_STABLE_DIRTY,
	CT_REFAT,
	CT_AWAKE_BODIES,
	CT_BODY_COUNT,
	CT_SOLVED_CONTACTS,
	CT_SOLVED_POINTS,
	CT_BB_POINTS,
	CT_AWAKE_ISLANDS,
	CT_INTEGRATED,
	CT_SEG_COUNT,
	CT_COLORED_ISLANDS,
	CT_SLOTS,
	CT_PAR_ENTRIES,
	CT_JOINT_ISLANDS,
	CT_GIANT_ISLANDS,
	CT_COLOR_ROUNDS,
	CT_GIANT_ENTRIES, // summed over substeps
	CT_GIANT_POINTS,  // summed over substeps
	CT_SUB_REBUILDS,  // batched: sub-worlds rebuilding their stable tier
	CT_KEPT,          // batched: stable entries kept from the last build
	CT_PLACED,        // batched: contacts after placement
	CT_DIFF_CONTACTS, // recorded solver contacts
	CT_GIANT_BB,      // box-box entries, summed over substeps
	CT_GIANT_BB_POINTS, // their points, summed over substeps
	CT_RESPAWN_REMOVED, // between steps: contacts touching the bodies of a bulk despawn
	CT_GRID_ABORT,      // a grid barrier's wait exceeded its bound (the step failed)
	CT_COUNT_
};

constexpr int kCountersSize = 64;
// The pinned host mirror holds, after the counters block, the first
// kOverHead oversized body ids of each tier: a tier build downloads them
// with its counter readback (more oversized bodies take a second copy).
constexpr int kOverHead = 256;
constexpr int kHostCountersWords = kCountersSize + 2 * kOverHead;

// island mode bits
constexpr uint32_t IM_AWAKE = 1u;
constexpr uint32_t IM_COLORED = 2u;
constexpr uint32_t IM_JOINT = 4u;
// awake islands without joints: solved together by the cooperative grid
constexpr uint32_t IM_GRID = 8u;

// Device island lists handed from island building to the island solve.
struct IslandLists
{
	const uint32_t* bodyList;     // awake island bodies, by island, ascending
	const uint32_t* bodyStart;    // per island, exclusive (I + 1)
	const uint32_t* contactList;  // solver contacts, by island, ascending
	const uint32_t* contactStart; // per island, exclusive (I + 1)
	const uint64_t* entryIsland;  // island of each contactList entry
	const uint32_t* mode;         // per island mode bits
	const int32_t* islandOf;      // per body
	const int32_t* jointIsland;   // per joint: awake island or -1
	int32_t I, NB, NC;
	int bitsI; // bits of an island id
};

// ---- k_broad.hip ----
void launch_refresh(World& w, float dt, bool rebuildStable);
// inStableTier := isStable for every slot (after a stable-tier build)
void launch_stable_migrate(World& w);
// Builds one tier (0 active, 1 stable); returns false on allocation failure.
bool build_tier(World& w, int tier);
bool enumerate_and_create(World& w);
bool remove_separated(World& w);
void launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
// Up to four fills of 32-bit words, one launch.
struct FillSegs
{
	uint32_t* p[4];
	uint32_t v[4];
	int64_t n[4];
	int count = 0;
	void add(uint32_t* ptr, uint32_t value, int64_t words)
	{
		p[count] = ptr;
		v[count] = value;
		n[count] = words;
		++count;
	}
};
void launch_fill_multi(const FillSegs& F, hipStream_t s);
bool apply_moves(World& w, const int2* dMoves, int32_t count);
void launch_material_refresh(World& w, const uint8_t* dFlags);
// applies and clears the host's list of bodies to wake
bool launch_wake_list(World& w);
// starts copying (body a, body b, flags) of every contact to pinned memory
bool prefetch_sidecar(World& w);
void launch_upload_scatter(World& w, int32_t count, const int32_t* dIdx, const float4* stage, const uint32_t* stageFlags,
						   const uint4* stageFilter, const uint8_t* stageBp);

// ---- batched sub-worlds (k_broad.hip) ----
// host-side sub-world state to the device (body sub-worlds, API dirt)
bool upload_batch_state(World& w);
// per sub-world stable-tier decision of this step (src/world.cpp:170, 212-227)
void launch_stable_plan(World& w);
// stable tier of the sub-worlds that rebuild it; the others keep theirs
bool build_stable_batched(World& w);
// groups the step's new pairs (candidate order) by sub-world
bool stage_new_contacts(World& w, const uint64_t* keys, int32_t K);
// separated-pair removal and new-contact placement per sub-world, each sub-world
// as src/world.cpp:1705-1735 and :1540-1568 would lay out a world of its own
bool remove_and_place(World& w);
bool batch_destroy_contacts(World& w, int32_t body, std::vector<int4>& events);

// ---- k_diff.hip ----
// records the step's solve inputs when differentiable recording is on
bool diff_record_step(World& w, const IslandLists& L, float dt, int substeps);
// records the manifold each substep's prepare left and the contacts' schedule
// words (after the solve records hold the substep's packet forms)
bool diff_record_substep(World& w, int sub, const uint32_t* atOf, const int4* rec);
bool diff_is_recording(const World& w);
// true (and the error set) when the recording was refused
bool diff_refused(const World& w);

// ---- bulk despawn (k_broad.hip) ----
// The contacts touching the bodies of a bulk despawn, as (body a, body b,
// flags, position) in ascending position: destroyed holds (slot, rank) of
// each destroyed body, rank = its destroy's place in the call.
bool gather_despawn_contacts(World& w, const std::vector<int2>& destroyed, std::vector<int4>& out);

// ---- k_query.hip ----
// sorted, repeat-free (query << 32 | body) overlap pairs of the query boxes
bool overlap_query(World& w, const m3d_aabb* boxes, int32_t count, std::vector<uint64_t>& pairs);

// ---- k_narrow.hip ----
bool narrowphase(World& w, float dt);
// issues the download of the narrowphase's transitions (counters read back)
bool queue_transitions(World& w);

// ---- k_solve.hip ----
bool islands_and_solve(World& w, float dt, int substeps);
void finish_step(World& w, float dt);

// ---- k_isolve.hip ----
bool solve_islands(World& w, const IslandLists& L, float dt, int substeps);

bool sync_counters(World& w);
// Sort histogram scratch and scan status words for scans and sorts of up
// to n elements.
bool ensure_prim_scratch(World& w, int64_t n);
bool push_host_changes(World& w);

} // namespace mx
