// m3d_world_step orchestration on the host: the order of stages follows
// src/world.cpp:1437-2167 (broadphase refresh, joint exclusion list, pair
// enumeration and creation, separated-pair removal, narrowphase and touch
// events, islands and solve, sleeping, kinematic bodies, force clear,
// profile). Every stage runs on the device; the host reads a small counter
// block between stages to size the next one.
#include "mx_steps.h"

#include <algorithm>
#include <chrono>
#include <unordered_map>
#include <string.h>

namespace mx
{

using Clock = std::chrono::steady_clock;

static float ms_between(Clock::time_point a, Clock::time_point b)
{
	return std::chrono::duration<float, std::milli>(b - a).count();
}

// Levels of the joints' serial order (index order, src/xpbd.cpp:913-969):
// a joint's level is one more than the highest level of the earlier joints
// on one of its dynamic bodies (only those are written), so the joints of
// one level touch disjoint bodies and running the levels in turn performs
// every body's joint updates in the serial order.
static bool joint_levels(World& w)
{
	w.jointLevelStart.clear();
	std::vector<int32_t> level(w.joints.size(), -1);
	std::unordered_map<int32_t, int32_t> last; // dynamic body -> level of its latest joint
	int32_t top = -1;
	for (size_t j = 0; j < w.joints.size(); ++j)
	{
		const DJoint& jt = w.joints[j];
		if (!jt.inUse)
			continue;
		const int32_t ends[2] = { jt.bodyA, jt.bodyB };
		int32_t lv = 0;
		for (int32_t b : ends)
		{
			auto it = w.bodies[b].type == M3D_BODY_DYNAMIC ? last.find(b) : last.end();
			if (it != last.end())
				lv = std::max(lv, it->second + 1);
		}
		for (int32_t b : ends)
			if (w.bodies[b].type == M3D_BODY_DYNAMIC)
				last[b] = lv;
		level[j] = lv;
		top = std::max(top, lv);
	}
	if (top < 0)
		return true;
	w.jointLevelStart.assign((size_t)top + 2, 0);
	for (int32_t lv : level)
		if (lv >= 0)
			++w.jointLevelStart[(size_t)lv + 1];
	for (int32_t l = 0; l <= top; ++l)
		w.jointLevelStart[(size_t)l + 1] += w.jointLevelStart[(size_t)l];
	std::vector<int32_t> order((size_t)w.jointLevelStart.back());
	std::vector<int32_t> fill(w.jointLevelStart.begin(), w.jointLevelStart.end() - 1);
	for (size_t j = 0; j < level.size(); ++j)
		if (level[j] >= 0)
			order[(size_t)fill[(size_t)level[j]]++] = (int32_t)j;
	return w.jointOrder.ensure(order.size()) &&
		   hip_ok(hipMemcpy(w.jointOrder.p, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice),
				  "joint levels");
}

static bool upload_joints(World& w)
{
	if (!w.jointsDirty)
		return true;
	int32_t nj = (int32_t)w.joints.size();
	w.djointCount = nj;
	if (nj > 0)
	{
		if (!w.djoints.ensure(nj))
			return false;
		if (!hip_ok(hipMemcpyAsync(w.djoints.p, w.joints.data(), nj * sizeof(DJoint), hipMemcpyHostToDevice, w.stream),
					"joint upload"))
			return false;
	}
	if (!joint_levels(w))
		return false;
	w.jointsDirty = false;
	return true;
}

// Joint lambdas/compliance are transient device state; bring them back so
// the host table (the upload source) stays current.
static bool download_joints(World& w)
{
	if (w.djointCount == 0)
		return true;
	return hip_ok(hipMemcpyAsync(w.joints.data(), w.djoints.p, w.djointCount * sizeof(DJoint), hipMemcpyDeviceToHost,
								 w.stream),
				  "joint download");
}

static bool upload_jointed_pairs(World& w)
{
	if (!w.jointedPairsDirty)
		return true;
	std::vector<uint64_t> keys;
	for (const DJoint& j : w.joints)
	{
		if (j.inUse)
		{
			int32_t lo = j.bodyA < j.bodyB ? j.bodyA : j.bodyB;
			int32_t hi = j.bodyA < j.bodyB ? j.bodyB : j.bodyA;
			keys.push_back(contact_key(lo, hi));
		}
	}
	std::sort(keys.begin(), keys.end());
	w.jointedCount = (int32_t)keys.size();
	if (!keys.empty())
	{
		if (!w.jointedKeys.ensure(keys.size()))
			return false;
		if (!hip_ok(hipMemcpy(w.jointedKeys.p, keys.data(), keys.size() * sizeof(uint64_t), hipMemcpyHostToDevice),
					"jointed pairs"))
			return false;
	}
	w.jointedPairsDirty = false;
	return true;
}

static bool apply_material_refresh(World& w)
{
	if (!w.anyMaterialRefresh)
		return true;
	int32_t n = w.deviceBodyCount;
	if (!w.u8a.ensure(n))
		return false;
	w.materialRefresh.resize(n, 0);
	hipMemcpyAsync(w.u8a.p, w.materialRefresh.data(), n, hipMemcpyHostToDevice, w.stream);
	launch_material_refresh(w, w.u8a.p);
	std::fill(w.materialRefresh.begin(), w.materialRefresh.end(), 0);
	w.anyMaterialRefresh = false;
	return hip_ok(hipStreamSynchronize(w.stream), "material refresh");
}

// Host changes made since the last step reach the device here.
bool push_host_changes(World& w)
{
	// wakes first: uploaded dirty records are authoritative and come after
	return apply_pending_permutation(w) && launch_wake_list(w) && upload_dirty(w) && upload_batch_state(w) &&
		   upload_joints(w) && apply_material_refresh(w);
}

// RefreshBroadphase; src/world.cpp:156-229
bool refresh_broadphase(World& w, float dt)
{
	if (w.subCount > 0)
	{
		// per sub-world stable-tier timing; the active tier is rebuilt for all
		launch_stable_plan(w);
		launch_refresh(w, dt, false);
		// the tier build reads the counters back after the plan (with no
		// bodies it launches nothing and reads nothing)
		if (!build_tier(w, 0) || (w.deviceBodyCount == 0 && !sync_counters(w)))
			return false;
		if (w.hostCounters[CT_SUB_REBUILDS] > 0)
		{
			build_stable_batched(w);
			launch_stable_migrate(w);
		}
		w.gridDirty = false;
		w.queryRefreshed = dt == 0.0f;
		w.gridBodies = w.deviceBodyCount;
		return true;
	}
	bool rebuildStable = w.stableForce || (w.stableDirty && w.stepsSinceStableBuild >= 4);
	launch_refresh(w, dt, rebuildStable);
	if (!build_tier(w, 0))
		return false;
	if (rebuildStable)
	{
		if (!build_tier(w, 1))
			return false;
		// inStableTier = isStable for every slot
		w.stableDirty = false;
		w.stableForce = false;
		w.stepsSinceStableBuild = 0;
		launch_stable_migrate(w);
	}
	else
	{
		++w.stepsSinceStableBuild;
	}
	w.gridDirty = false;
	w.queryRefreshed = dt == 0.0f;
	w.gridBodies = w.deviceBodyCount;
	return true;
}

static bool world_step_stages(World& w, float dt, int substeps);

void world_step(World& w, float dt, int substeps)
{
	if (dt <= 0.0f)
		return;
	// sorts are timed (M3DX_K_SORT) while the world's kernel timing is on
	SortTimer timer = { &w, [](void* c, double bytes) { static_cast<World*>(c)->mark_begin(M3DX_K_SORT, bytes); },
						[](void* c) { static_cast<World*>(c)->mark_end(); } };
	set_sort_timer(w.timing ? &timer : nullptr);
	bool ok = world_step_stages(w, dt, substeps < 1 ? 1 : substeps);
	set_sort_timer(nullptr);
	if (!ok)
		++w.stepFailures;
}

// The step's stages; false when one stopped on an error.
static bool world_step_stages(World& w, float dt, int substeps)
{
	auto t0 = Clock::now();
	w.beginEvents.clear();
	w.endEvents.clear();
	w.pendingTrans = 0;
	w.transInFlight = false;
	memset(&w.stats, 0, sizeof(w.stats));
	// a sidecar downloaded before this step (and not used by a destroy since)
	// no longer matches the device's contacts once the step runs
	w.sidecarPrefetched = false;
	if (!push_host_changes(w))
		return false;
	hipMemsetAsync(w.counters.p, 0, kCountersSize * sizeof(uint32_t), w.stream);

	if (!refresh_broadphase(w, dt))
		return false;
	if (!upload_jointed_pairs(w) || !enumerate_and_create(w))
		return false;
	if (!(w.subCount > 0 ? remove_and_place(w) : remove_separated(w)))
		return false;
	auto tBroad = Clock::now();
	if (!narrowphase(w, dt))
		return false;
	auto tNarrow = Clock::now();
	if (!islands_and_solve(w, dt, substeps))
		return false;
	finish_step(w, dt);
	if (!flush_transitions(w) || !download_joints(w) || !sync_counters(w))
		return false;
	const uint32_t* hc = w.hostCounters;
	if (hc[CT_GRID_ABORT] != 0u)
	{
		// the grid was not co-resident: this step's solve is void
		w.coopLaunch = true;
		set_error("a grid barrier wait exceeded %u ticks; the world launches its grid kernels cooperatively from now on",
				  w.barTimeout);
		return false;
	}
	w.tuner_step_end((int32_t)hc[CT_SLOTS]);
	if (w.timing)
	{
		// algorithmic bytes of the timed kernels this step, each live field
		// read or written once per pass (entries and points summed over the
		// substeps):
		// - normals, per contact and iteration 160 B + 68 B per point;
		// - stiction, 16 B per record read + per box-box contact 192 B (two
		//   body poses read and written) + 72 B per box-box point;
		// - velocity, per contact 472 B + 112 B per point;
		// - prepare, per contact 388 B (record, bodies, material, normal,
		//   pose cache, both bodies' pose, shape, inertia) + 272 B per point
		//   (anchors, slip and arms read; rows written to the contact and
		//   the schedule-ordered copy) + 64 B per box-box point (friction
		//   anchors copied);
		// - bodies, 184 B per integrated body per substep (integration and
		//   reconstruction together);
		// - ordering, 48 B per solver contact and fixpoint round;
		// - pairs, 32 B per grid entry + 100 B per candidate pair;
		// - narrowphase, 640 B per regenerated manifold, 120 B per cached
		//   contact;
		// - sorts, counted per launch (32 B per element and pass).
		double ge = (double)hc[CT_GIANT_ENTRIES], gp = (double)hc[CT_GIANT_POINTS];
		double gbb = (double)hc[CT_GIANT_BB], gbbp = (double)hc[CT_GIANT_BB_POINTS];
		int iters = w.def.positionIterations > 0 ? w.def.positionIterations : 2;
		double regen = (double)hc[CT_REGEN];
		double stepBytes[M3DX_K_COUNT] = {};
		stepBytes[M3DX_K_GIANT_NORMALS] = iters * (160.0 * ge + 68.0 * gp);
		stepBytes[M3DX_K_GIANT_FRICTION] = 16.0 * ge + 192.0 * gbb + 72.0 * gbbp;
		stepBytes[M3DX_K_GIANT_VELOCITY] = 472.0 * ge + 112.0 * gp;
		stepBytes[M3DX_K_GIANT_PREPARE] = 388.0 * ge + 272.0 * gp + 64.0 * gbbp;
		stepBytes[M3DX_K_GIANT_BODIES] = 184.0 * substeps * (double)hc[CT_INTEGRATED];
		stepBytes[M3DX_K_GIANT_ORDER] = 48.0 * (double)w.stats.solvedContacts * (double)hc[CT_COLOR_ROUNDS];
		stepBytes[M3DX_K_PAIRS] = 32.0 * (double)(w.activeCount + w.stableCount) + 100.0 * (double)hc[CT_CANDIDATES];
		stepBytes[M3DX_K_NARROW] = 640.0 * regen + 120.0 * std::max(0.0, (double)w.contactCount - regen);
		w.collect_marks();
		for (int k = 0; k < M3DX_K_COUNT; ++k)
			w.ktime[k].bytes += stepBytes[k];
	}
	if (hc[CT_STABLE_DIRTY])
		w.stableDirty = true;
	w.anyForcesApplied = false;
	auto tEnd = Clock::now();

	w.profile.broadphaseMs = ms_between(t0, tBroad);
	w.profile.narrowphaseMs = ms_between(tBroad, tNarrow);
	w.profile.solverMs = ms_between(tNarrow, tEnd);
	w.profile.stepMs = ms_between(t0, tEnd);
	w.profile.bodyCount = (int32_t)hc[CT_BODY_COUNT];
	w.profile.awakeBodyCount = (int32_t)hc[CT_AWAKE_BODIES];
	w.profile.contactCount = w.contactCount;
	w.profile.touchingContactCount = w.lastHadIslands ? (int32_t)hc[CT_TOUCHING] : 0;
	w.profile.islandCount = w.lastHadIslands ? w.lastIslandCount : 0;

	w.stats.bodyCapacity = w.deviceBodyCount;
	w.stats.contacts = w.contactCount;
	w.stats.regeneratedManifolds = (int32_t)hc[CT_REGEN];
	w.stats.activeEntries = (int32_t)w.activeCount;
	w.stats.stableEntries = (int32_t)w.stableCount;
	w.stats.refattedBodies = (int32_t)hc[CT_REFAT];
	w.stats.islands = w.profile.islandCount;
	w.stats.sweeps = (int32_t)hc[CT_SLOTS];
	w.stats.colorRounds = (int32_t)hc[CT_COLOR_ROUNDS];

	w.hostStale = true;
	w.sidecarValid = false;
	// a caller that destroyed bodies before this step likely does again
	// before the next: start the contact sidecar download now
	if (w.destroyedSinceStep)
		(void)prefetch_sidecar(w);
	w.destroyedSinceStep = false;
	return true;
}

} // namespace mx
