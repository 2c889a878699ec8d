// The m3d_* C API over the device backend, plus the m3dx_* extension.
//
// Semantics of every entry point follow the reference implementation cited
// beside it (src/world.cpp): defaults, id validation, LIFO slot reuse,
// generation bumps, wake rules, error returns. Body state edits go to the
// host mirror and are uploaded at the next step; reads refresh the mirror
// from the device once per step.
#include "mx_steps.h"

#include <algorithm>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>

namespace mx
{

static thread_local char g_error[512] = "";
static int g_device = 0;

void set_error(const char* fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_error, sizeof(g_error), fmt, ap);
	va_end(ap);
	fprintf(stderr, "melon3d_amd: %s\n", g_error);
}

bool hip_ok(hipError_t e, const char* what)
{
	if (e == hipSuccess)
		return true;
	set_error("%s: %s", what, hipGetErrorString(e));
	return false;
}

World::World()
{
	// M3D_STEP_GRAPHS=0 launches a batch's substep loop directly (profilers
	// that intercept graph launches)
	const char* g = getenv("M3D_STEP_GRAPHS");
	stepGraphs = !(g != nullptr && g[0] == '0');
}

World::~World()
{
	diff_release(*this);
	BodyColumns& b = bc;
	DBuf<float4>* f4[] = { &b.pos, &b.rot, &b.vel, &b.angVel, &b.prevPos, &b.prevRot, &b.prevVel, &b.prevAngVel,
						   &b.invIL, &b.iw0, &b.iw1, &b.iw2, &b.shape, &b.mat, &b.force, &b.torque, &b.fatLo, &b.fatHi,
						   &cc.mat, &cc.normal, &cc.cache, &cc.pA, &cc.pB,
						   &cc.pJA, &cc.pJB, &cc.pRA, &cc.pRB, &cc.pFA, &cc.pFB, &cc.pSlip };
	for (auto* p : f4)
		p->release();
	b.flags.release();
	b.filter.release();
	b.bp.release();
	b.sb.release();
	cc.bodies.release();
	cc.hint.release();
	{
		DBuf<float4>* c2[] = { &cc2.mat, &cc2.normal, &cc2.cache,
							   &cc2.pA, &cc2.pB, &cc2.pJA, &cc2.pJB, &cc2.pRA, &cc2.pRB, &cc2.pFA, &cc2.pFB,
							   &cc2.pSlip };
		for (auto* p : c2)
			p->release();
		cc2.bodies.release();
		cc2.hint.release();
	}
	bodySub.release();
	subStable.release();
	subDirtyUpload.release();
	subContactStart.release();
	subWork.release();
	newKeys.release();
	djoints.release();
	jointOrder.release();
	jointedKeys.release();
	activeKeys.release();
	stableKeys.release();
	stableKeysNext.release();
	stableBodiesNext.release();
	sortKeysAlt.release();
	activeBodies.release();
	stableBodies.release();
	sortValsAlt.release();
	dActiveOversized.release();
	dStableOversized.release();
	u32a.release();
	u32b.release();
	u32c.release();
	u32d.release();
	u64a.release();
	u64b.release();
	u64c.release();
	i32a.release();
	i32b.release();
	i32c.release();
	i32d.release();
	i32e.release();
	transList.release();
	transInfo.release();
	u8a.release();
	primScratch.release();
	pairRuns.release();
	boundParts.release();
	subBodyStartD.release();
	subEntryStart.release();
	overSegs.release();
	hashKeys.release();
	scanStatus.release();
	hashVals.release();
	islandOf.release();
	islandAwakeD.release();
	islandMode.release();
	s64a.release();
	s64b.release();
	s32a.release();
	s32b.release();
	s32c.release();
	sSlotStart.release();
	sColor.release();
	sLevel.release();
	sJointList.release();
	sCtr.release();
	sPacket.release();
	sRecG.release();
	sPosG.release();
	sAtOf.release();
	bEvFlag.release();
	bEvOff.release();
	bEvIdx.release();
	bEvOut.release();
	gridBar.release();
	tailOrder.release();
	sNrm.release();
	sBucket.release();
	sPred.release();
	sColored.release();
	sDirty.release();
	dWake.release();
	contactColor.release();
	counters.release();
	moves.release();
	eventIdx.release();
	if (hostCounters != nullptr)
		(void)hipHostFree(hostCounters);
	for (hipEvent_t e : eventPool)
		(void)hipEventDestroy(e);
	if (stepGraphExec != nullptr)
		(void)hipGraphExecDestroy(stepGraphExec);
	if (tuner.ev0 != nullptr)
		(void)hipEventDestroy(tuner.ev0);
	if (tuner.ev1 != nullptr)
		(void)hipEventDestroy(tuner.ev1);
	if (sidecarReady != nullptr)
		(void)hipEventDestroy(sidecarReady);
	if (pinnedSidecar != nullptr)
		(void)hipHostFree(pinnedSidecar);
	if (pinnedCsr != nullptr)
		(void)hipHostFree(pinnedCsr);
	dCsr.release();
	csrKeys.release();
	csrKeysAlt.release();
	csrVals.release();
	csrValsAlt.release();
	dSidecar.release();
	qExtra.release();
	qFlag.release();
	qOffs.release();
	qUFlag.release();
	qUOffs.release();
	qPairs.release();
	qPairsAlt.release();
	qVals.release();
	qValsAlt.release();
	respRank.release();
	respList.release();
	respOut.release();
	if (pinnedResp != nullptr)
		(void)hipHostFree(pinnedResp);
	if (uploadDone != nullptr)
		(void)hipEventDestroy(uploadDone);
	if (pinnedUpload != nullptr)
		(void)hipHostFree(pinnedUpload);
	if (pinnedRemoval != nullptr)
		(void)hipHostFree(pinnedRemoval);
	if (pinnedSegs != nullptr)
		(void)hipHostFree(pinnedSegs);
	if (transReady != nullptr)
		(void)hipEventDestroy(transReady);
	if (pinnedTrans != nullptr)
		(void)hipHostFree(pinnedTrans);
	if (stream != nullptr)
		(void)hipStreamDestroy(stream);
}

// ------------------------------------------------------------- kernel timing

void World::mark_begin(int k, double bytes)
{
	if (!timing)
		return;
	while (eventPool.size() < eventsUsed + 2)
	{
		hipEvent_t e = nullptr;
		if (hipEventCreate(&e) != hipSuccess)
			return;
		eventPool.push_back(e);
	}
	KernelMark m;
	m.kernel = k;
	m.begin = eventPool[eventsUsed++];
	m.end = eventPool[eventsUsed++];
	m.bytes = bytes;
	(void)hipEventRecord(m.begin, stream);
	marks.push_back(m);
}

void World::mark_end()
{
	if (!timing || marks.empty())
		return;
	(void)hipEventRecord(marks.back().end, stream);
}

void World::collect_marks()
{
	for (const KernelMark& m : marks)
	{
		float ms = 0.0f;
		if (hipEventElapsedTime(&ms, m.begin, m.end) == hipSuccess)
		{
			m3dx_kernel_time& t = ktime[m.kernel];
			t.launches += 1;
			t.ms += ms;
			t.bytes += m.bytes;
		}
	}
	marks.clear();
	eventsUsed = 0;
}

// ------------------------------------------------------------- pass tuner

constexpr int World::kTunerCands[World::PassTuner::kCands];

int World::tuner_pick(int32_t entries, int coopMax, bool& measure)
{
	measure = false;
	PassTuner& T = tuner;
	// more workgroups than one lane per entry only add barrier arrivals
	if (T.coopLimit > 0)
		coopMax = std::min(coopMax, T.coopLimit);
	int need = (int)std::min<int64_t>(((int64_t)entries + 255) / 256, (int64_t)std::max(coopMax, 1));
	need = std::max(need, 1);
	if (T.forcedBlocks > 0)
		return std::min(T.forcedBlocks, need);
	if (need <= kTunerCands[0])
		return need;
	int c = 0;
	while (c + 1 < PassTuner::kClasses && ((int64_t)1 << (c + 1)) <= (int64_t)entries)
		++c;
	PassTuner::Class& K = T.cls[c];
	if (K.chosen > 0)
		return std::min(K.chosen, need);
	if (T.ev0 == nullptr && hipEventCreate(&T.ev0) != hipSuccess)
		T.ev0 = nullptr;
	if (T.ev1 == nullptr && hipEventCreate(&T.ev1) != hipSuccess)
		T.ev1 = nullptr;
	if (T.ev0 == nullptr || T.ev1 == nullptr)
		return std::min(kTunerCands[1], need);
	int cand = K.trials % PassTuner::kCands;
	T.pendingClass = c;
	T.pendingCand = cand;
	measure = true;
	return std::min(kTunerCands[cand], need);
}

void World::tuner_step_end(int32_t slots)
{
	PassTuner& T = tuner;
	if (T.pendingClass < 0)
		return;
	float ms = 0.0f;
	if (hipEventElapsedTime(&ms, T.ev0, T.ev1) == hipSuccess)
	{
		PassTuner::Class& K = T.cls[T.pendingClass];
		// a candidate the runtime refused (and ran with fewer) is never chosen
		float per = T.pendingFailed ? 3.0e38f : ms / (float)std::max(slots, 1);
		float& best = K.perSlot[T.pendingCand];
		best = K.trials < PassTuner::kCands ? per : std::min(best, per);
		if (++K.trials >= PassTuner::kCands * PassTuner::kRounds)
		{
			int arg = 0;
			for (int k = 1; k < PassTuner::kCands; ++k)
				if (K.perSlot[k] < K.perSlot[arg])
					arg = k;
			K.chosen = kTunerCands[arg];
		}
	}
	T.pendingClass = -1;
	T.pendingCand = -1;
	T.pendingFailed = false;
}

// ------------------------------------------------------------- transfers

template <class T>
static bool grow_keep(DBuf<T>& b, size_t n, size_t keep)
{
	return b.ensure(n, keep);
}

bool ensure_device_bodies(World& w, int32_t n)
{
	if (n <= w.deviceBodyCap)
		return true;
	size_t keep = (size_t)w.deviceBodyCount;
	size_t cap = std::max<size_t>((size_t)n, (size_t)w.deviceBodyCap * 3 / 2 + 64);
	BodyColumns& b = w.bc;
	DBuf<float4>* f4[] = { &b.pos, &b.rot, &b.vel, &b.angVel, &b.prevPos, &b.prevRot, &b.prevVel, &b.prevAngVel,
						   &b.invIL, &b.iw0, &b.iw1, &b.iw2, &b.shape, &b.mat, &b.force, &b.torque, &b.fatLo, &b.fatHi };
	for (auto* p : f4)
		if (!grow_keep(*p, cap, keep))
			return false;
	if (!grow_keep(b.flags, cap, keep) || !grow_keep(b.filter, cap, keep) || !grow_keep(b.bp, cap, keep) ||
		!grow_keep(b.sb, cap, keep))
		return false;
	w.deviceBodyCap = (int32_t)b.pos.cap;
	return true;
}

template <class T>
static bool realloc_exact(DBuf<T>& b, size_t cap, size_t keep)
{
	T* np = nullptr;
	if (!hip_ok(hipMalloc((void**)&np, cap * sizeof(T)), "contact columns"))
		return false;
	if (b.p != nullptr)
	{
		if (keep > 0)
			(void)hipMemcpy(np, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice);
		(void)hipFree(b.p);
	}
	b.p = np;
	b.cap = cap;
	return true;
}

// Contact point columns keep a contact's four points together ([4c + j]); growing
// copies the kept contacts' points.
static bool grow_points(DBuf<float4>& col, size_t oldCap, size_t newCap, size_t keep)
{
	float4* np = nullptr;
	if (!hip_ok(hipMalloc((void**)&np, 4 * newCap * sizeof(float4)), "contact point columns"))
		return false;
	if (col.p != nullptr)
	{
		// a contact's four points are contiguous (point j of contact c at 4c + j)
		if (keep > 0)
			(void)hipMemcpy(np, col.p, 4 * keep * sizeof(float4), hipMemcpyDeviceToDevice);
		(void)hipFree(col.p);
	}
	col.p = np;
	col.cap = 4 * newCap;
	return true;
}

bool ensure_device_contacts(World& w, int32_t n)
{
	return ensure_contact_columns(w, w.cc, w.contactCap, n, (size_t)w.contactCount);
}

bool ensure_contact_columns(World& w, ContactColumns& c, int32_t& capRef, int32_t n, size_t keep)
{
	if (n <= capRef)
		return true;
	(void)hipStreamSynchronize(w.stream);
	size_t oldCap = (size_t)capRef;
	size_t cap = std::max<size_t>((size_t)n, oldCap * 3 / 2 + 256);
	DBuf<float4>* cols[] = { &c.mat, &c.normal };
	for (auto* p : cols)
		if (!realloc_exact(*p, cap, keep))
			return false;
	if (!realloc_exact(c.bodies, cap, keep) || !realloc_exact(c.hint, cap, keep))
		return false;
	DBuf<float4>* pts[] = { &c.pA, &c.pB, &c.pJA, &c.pJB, &c.pRA, &c.pRB, &c.pFA, &c.pFB, &c.pSlip, &c.cache };
	for (auto* p : pts)
		if (!grow_points(*p, oldCap, cap, keep))
			return false;
	capRef = (int32_t)cap;
	return true;
}

static uint32_t pack_flags(const HostBody& b)
{
	uint32_t f = (uint32_t)b.type & BF_TYPE_MASK;
	f |= ((uint32_t)b.shape.type & 3u) << BF_SHAPE_SHIFT;
	if (b.isSensor)
		f |= BF_SENSOR;
	if (b.enableSleep)
		f |= BF_ENABLE_SLEEP;
	if (b.isAwake)
		f |= BF_AWAKE;
	if (b.inUse)
		f |= BF_IN_USE;
	return f;
}

static void pack_columns(const HostBody& b, float4* c /* 18 */)
{
	c[0] = v3_f4(b.position, b.invMass);
	c[1] = q_f4(b.rotation);
	c[2] = v3_f4(b.linearVelocity, b.gravityScale);
	c[3] = v3_f4(b.angularVelocity, b.mass);
	c[4] = v3_f4(b.prevPosition, 0.0f);
	c[5] = q_f4(b.prevRotation);
	c[6] = v3_f4(b.prevLinearVelocity, 0.0f);
	c[7] = v3_f4(b.prevAngularVelocity, 0.0f);
	c[8] = v3_f4(b.invInertiaLocal, b.linearDamping);
	c[9] = v3_f4(b.invInertiaWorld.cx, b.angularDamping);
	c[10] = v3_f4(b.invInertiaWorld.cy, b.sleepTime);
	c[11] = v3_f4(b.invInertiaWorld.cz, b.smoothedEnergy);
	c[12] = v3_f4(b.shape.halfExtents, b.shape.radius);
	c[13] = make_float4(b.shape.halfHeight, b.friction, b.restitution, b.rollingResistance);
	c[14] = v3_f4(b.force, 0.0f);
	c[15] = v3_f4(b.torque, 0.0f);
	c[16] = v3_f4(b.fat.lowerBound, 0.0f);
	c[17] = v3_f4(b.fat.upperBound, 0.0f);
}

bool upload_dirty(World& w)
{
	int32_t n = (int32_t)w.bodies.size();
	if (!ensure_device_bodies(w, n))
		return false;
	if (n > w.deviceBodyCount)
	{
		// new slots are always dirty (created since the last step)
		w.deviceBodyCount = n;
	}
	int32_t k = (int32_t)w.dirtyList.size();
	if (k == 0)
		return true;
	// one block, laid out as on the device: 18 float4 columns of k, then
	// filters, flags, slots and tier bits; staged in pinned memory so the
	// copy and the scatter run behind the host (the stream orders every
	// later use of the device block; the pinned block is reused only after
	// its last copy completed)
	size_t colBytes = (size_t)k * 18 * sizeof(float4);
	size_t bytes = colBytes + (size_t)k * (sizeof(uint4) + sizeof(uint32_t) + sizeof(int32_t) + 1);
	if (w.uploadDone != nullptr && !hip_ok(hipEventSynchronize(w.uploadDone), "upload staging"))
		return false;
	if (w.pinnedUploadCap < bytes)
	{
		if (w.pinnedUpload != nullptr)
			(void)hipHostFree(w.pinnedUpload);
		w.pinnedUpload = nullptr;
		w.pinnedUploadCap = bytes * 3 / 2 + 4096;
		if (!hip_ok(hipHostMalloc((void**)&w.pinnedUpload, w.pinnedUploadCap, hipHostMallocDefault), "upload staging"))
		{
			w.pinnedUploadCap = 0;
			return false;
		}
	}
	if (w.uploadDone == nullptr && !hip_ok(hipEventCreateWithFlags(&w.uploadDone, hipEventDisableTiming), "upload event"))
		return false;
	char* hb = w.pinnedUpload;
	float4* stage = (float4*)hb;
	uint4* sfil = (uint4*)(hb + colBytes);
	uint32_t* sf = (uint32_t*)(sfil + k);
	int32_t* sidx = (int32_t*)(sf + k);
	uint8_t* sbp = (uint8_t*)(sidx + k);
	float4 cols[18];
	for (int32_t t = 0; t < k; ++t)
	{
		const HostBody& b = w.bodies[w.dirtyList[t]];
		pack_columns(b, cols);
		for (int c = 0; c < 18; ++c)
			stage[(size_t)c * k + t] = cols[c];
		sf[t] = pack_flags(b);
		sfil[t] = make_uint4(b.filter.categoryBits, b.filter.maskBits, (uint32_t)b.filter.groupIndex, 0u);
		sidx[t] = w.dirtyList[t];
		sbp[t] = b.bp;
	}
	if (!w.u64a.ensure(bytes / 8 + 1))
		return false;
	char* base = (char*)w.u64a.p;
	float4* dStage = (float4*)base;
	uint4* dFil = (uint4*)(base + colBytes);
	uint32_t* dFlags = (uint32_t*)(dFil + k);
	int32_t* dIdx = (int32_t*)(dFlags + k);
	uint8_t* dBp = (uint8_t*)(dIdx + k);
	hipStream_t s = w.stream;
	if (!hip_ok(hipMemcpyAsync(base, hb, bytes, hipMemcpyHostToDevice, s), "upload"))
		return false;
	launch_upload_scatter(w, k, dIdx, dStage, dFlags, dFil, dBp);
	if (!hip_ok(hipEventRecord(w.uploadDone, s), "upload event"))
		return false;
	for (int32_t i : w.dirtyList)
		w.dirtyFlag[i] = 0;
	w.dirtyList.clear();
	return true;
}

bool download_bodies(World& w)
{
	int32_t n = std::min<int32_t>(w.deviceBodyCount, (int32_t)w.bodies.size());
	if (n == 0)
	{
		w.hostStale = false;
		return true;
	}
	BodyColumns& b = w.bc;
	DBuf<float4>* f4[] = { &b.pos, &b.rot, &b.vel, &b.angVel, &b.prevPos, &b.prevRot, &b.prevVel, &b.prevAngVel,
						   &b.invIL, &b.iw0, &b.iw1, &b.iw2, &b.shape, &b.mat, &b.force, &b.torque, &b.fatLo, &b.fatHi };
	std::vector<float4> host((size_t)n * 18);
	std::vector<uint32_t> flags(n);
	std::vector<uint8_t> bp(n);
	hipStream_t s = w.stream;
	for (int c = 0; c < 18; ++c)
		if (!hip_ok(hipMemcpyAsync(host.data() + (size_t)c * n, f4[c]->p, n * sizeof(float4), hipMemcpyDeviceToHost, s),
					"download"))
			return false;
	if (!hip_ok(hipMemcpyAsync(flags.data(), b.flags.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "download") ||
		!hip_ok(hipMemcpyAsync(bp.data(), b.bp.p, n, hipMemcpyDeviceToHost, s), "download") ||
		!hip_ok(hipStreamSynchronize(s), "download"))
		return false;
	for (int32_t i = 0; i < n; ++i)
	{
		if (!w.dirtyFlag.empty() && i < (int32_t)w.dirtyFlag.size() && w.dirtyFlag[i])
			continue; // host edit pending: the mirror is authoritative
		HostBody& hb = w.bodies[i];
		auto col = [&](int c) { return host[(size_t)c * n + i]; };
		float4 c0 = col(0), c2 = col(2), c3 = col(3), c8 = col(8), c9 = col(9), c10 = col(10), c11 = col(11);
		hb.position = f4_v3(c0);
		hb.rotation = f4_q(col(1));
		hb.linearVelocity = f4_v3(c2);
		hb.angularVelocity = f4_v3(c3);
		hb.prevPosition = f4_v3(col(4));
		hb.prevRotation = f4_q(col(5));
		hb.prevLinearVelocity = f4_v3(col(6));
		hb.prevAngularVelocity = f4_v3(col(7));
		hb.invInertiaWorld.cx = f4_v3(c9);
		hb.invInertiaWorld.cy = f4_v3(c10);
		hb.invInertiaWorld.cz = f4_v3(c11);
		hb.sleepTime = c10.w;
		hb.smoothedEnergy = c11.w;
		hb.force = f4_v3(col(14));
		hb.torque = f4_v3(col(15));
		hb.fat.lowerBound = f4_v3(col(16));
		hb.fat.upperBound = f4_v3(col(17));
		hb.isAwake = (flags[i] & BF_AWAKE) != 0;
		hb.bp = bp[i] & (BP_HAS_FAT | BP_IN_STABLE_TIER | BP_IN_STABLE_CELLS);
		(void)c8;
	}
	w.hostStale = false;
	return true;
}

void ensure_host(World& w)
{
	if (w.hostStale)
	{
		launch_wake_list(w); // pending wakes belong to the state being read
		download_bodies(w);
	}
}

void mark_dirty(World& w, int32_t i)
{
	if ((int32_t)w.dirtyFlag.size() < (int32_t)w.bodies.size())
		w.dirtyFlag.resize(w.bodies.size(), 0);
	if (!w.dirtyFlag[i])
	{
		w.dirtyFlag[i] = 1;
		w.dirtyList.push_back(i);
	}
}

bool ensure_sidecar(World& w)
{
	if (w.sidecarValid)
		return true;
	int32_t C = w.contactCount;
	int32_t n = w.deviceBodyCount;
	w.hcCount = C;
	w.hcN = 0;
	w.hcMoved.clear();
	if (C > 0)
	{
		if (!(w.sidecarPrefetched && w.sidecarPrefetchCount == C && w.sidecarPrefetchN == n) && !prefetch_sidecar(w))
			return false;
		if (!hip_ok(hipEventSynchronize(w.sidecarReady), "sidecar"))
			return false;
		w.hcSide = w.pinnedSidecar;
		w.hcN = n;
		w.hcStart = w.pinnedCsr;
		w.hcList = w.pinnedCsr + n + 1;
		w.hcSlot = (int2*)(w.hcList + 2 * (size_t)C);
	}
	w.sidecarPrefetched = false;
	w.sidecarValid = true;
	w.pendingPermutation = false;
	return true;
}

bool apply_pending_permutation(World& w)
{
	if (!w.pendingPermutation)
		return true;
	int32_t newC = w.hcCount;
	std::vector<int2> mv;
	for (int32_t i : w.hcMoved)
	{
		// the contact now at i came from its original slot (always at or
		// beyond newC, so sources and destinations are disjoint)
		if (i < newC && w.hcSide[i].w != i)
		{
			mv.push_back(make_int2(i, w.hcSide[i].w));
			w.hcSide[i].w = i;
		}
	}
	if (!mv.empty())
	{
		if (!w.moves.ensure(mv.size()))
			return false;
		if (!hip_ok(hipMemcpyAsync(w.moves.p, mv.data(), mv.size() * sizeof(int2), hipMemcpyHostToDevice, w.stream),
					"permutation"))
			return false;
		apply_moves(w, w.moves.p, (int32_t)mv.size());
	}
	w.contactCount = newC;
	w.hcMoved.clear();
	w.pendingPermutation = false;
	return hip_ok(hipStreamSynchronize(w.stream), "permutation");
}

// stableDirty of the world holding body i (per sub-world when batched)
void set_stable_dirty(World& w, int32_t i)
{
	if (w.subCount > 0)
	{
		w.subDirtyHost[w.bodySubHost[i]] |= (uint8_t)SUB_DIRTY;
		w.anySubDirtyHost = true;
	}
	else
	{
		w.stableDirty = true;
	}
}

// WakeBodyInternal on the host mirror; src/world.cpp:121-140
void wake_body_host(World& w, int32_t i)
{
	HostBody& b = w.bodies[i];
	if (b.type == M3D_BODY_STATIC)
		return;
	if (!b.isAwake)
		set_stable_dirty(w, i);
	b.isAwake = true;
	b.sleepTime = 0.0f;
	b.smoothedEnergy = 4.0f * kSleepEnergyThresh;
	mark_dirty(w, i);
}

// mass properties; src/world.cpp:46-98
static void compute_mass(HostBody& b, const m3d_shape_def* def)
{
	float density = def->density > 0.0f ? def->density : 1000.0f;
	float mass = 0.0f;
	m3d_vec3 inertia = m3d_v3_zero();
	switch (def->type)
	{
		case M3D_SHAPE_SPHERE:
		{
			float r = def->radius;
			mass = density * (4.0f / 3.0f) * M3D_PI * r * r * r;
			float i = 0.4f * mass * r * r;
			inertia = m3d_v3(i, i, i);
			break;
		}
		case M3D_SHAPE_BOX:
		{
			m3d_vec3 e = def->halfExtents;
			mass = density * 8.0f * e.x * e.y * e.z;
			inertia = m3d_v3(mass / 3.0f * (e.y * e.y + e.z * e.z), mass / 3.0f * (e.x * e.x + e.z * e.z),
							 mass / 3.0f * (e.x * e.x + e.y * e.y));
			break;
		}
		case M3D_SHAPE_CAPSULE:
		{
			float r = def->radius;
			float h = def->halfHeight;
			float mc = density * M3D_PI * r * r * (2.0f * h);
			float ms = density * (4.0f / 3.0f) * M3D_PI * r * r * r;
			mass = mc + ms;
			float iy = 0.5f * mc * r * r + 0.4f * ms * r * r;
			float d = h + 0.375f * r;
			float ixz = mc * (0.25f * r * r + h * h / 3.0f) + (83.0f / 320.0f) * ms * r * r + ms * d * d;
			inertia = m3d_v3(ixz, iy, ixz);
			break;
		}
	}
	b.mass = mass;
	if (b.type == M3D_BODY_DYNAMIC && mass > 0.0f)
	{
		b.invMass = 1.0f / mass;
		b.invInertiaLocal = m3d_v3(inertia.x > 0.0f ? 1.0f / inertia.x : 0.0f, inertia.y > 0.0f ? 1.0f / inertia.y : 0.0f,
								   inertia.z > 0.0f ? 1.0f / inertia.z : 0.0f);
	}
	else
	{
		b.invMass = 0.0f;
		b.invInertiaLocal = m3d_v3_zero();
	}
}

} // namespace mx

using namespace mx;

// ------------------------------------------------------------ helpers

static HostBody* body_checked(World& w, m3d_body_id id)
{
	if (id.index < 0 || id.index >= (int32_t)w.bodies.size())
		return nullptr;
	HostBody& b = w.bodies[id.index];
	if (!b.inUse || b.generation != id.generation)
		return nullptr;
	return &b;
}

static m3d_body_id make_id(World& w, int32_t i) { return { i, w.bodies[i].generation }; }

static void update_inv_inertia(HostBody& b) { b.invInertiaWorld = inv_inertia_world(b.rotation, b.invInertiaLocal); }

static DJoint* joint_checked(World& w, m3d_joint_id id)
{
	if (id.index < 0 || id.index >= (int32_t)w.joints.size())
		return nullptr;
	DJoint& j = w.joints[id.index];
	if (!j.inUse || w.jointGen[id.index] != id.generation)
		return nullptr;
	return &j;
}

static WB wb_of(const HostBody& b)
{
	WB r;
	r.p = b.position;
	r.q = b.rotation;
	r.v = b.linearVelocity;
	r.w = b.angularVelocity;
	r.invMass = b.invMass;
	r.iw = b.invInertiaWorld;
	return r;
}

// ------------------------------------------------------------ C API

extern "C"
{

// defaults; src/world.cpp:400-462
m3d_world_def m3d_world_def_default(void)
{
	m3d_world_def d;
	d.gravity = m3d_v3(0.0f, -9.81f, 0.0f);
	d.workerCount = 1;
	d.positionIterations = 2;
	d.gridCellSize = 1.5f;
	d.enableSleep = true;
	return d;
}

m3d_body_def m3d_body_def_default(void)
{
	m3d_body_def d;
	d.type = M3D_BODY_STATIC;
	d.position = m3d_v3_zero();
	d.orientation = m3d_quat_identity();
	d.linearVelocity = m3d_v3_zero();
	d.angularVelocity = m3d_v3_zero();
	d.linearDamping = 0.0f;
	d.angularDamping = 0.05f;
	d.gravityScale = 1.0f;
	d.enableSleep = true;
	d.isAwake = true;
	d.userData = 0;
	return d;
}

m3d_shape_def m3d_shape_def_default(void)
{
	m3d_shape_def d;
	d.type = M3D_SHAPE_SPHERE;
	d.radius = 0.5f;
	d.halfHeight = 0.5f;
	d.halfExtents = m3d_v3(0.5f, 0.5f, 0.5f);
	d.density = 1000.0f;
	d.friction = 0.6f;
	d.restitution = 0.0f;
	d.rollingResistance = 0.02f;
	d.filter = m3d_filter_default();
	d.isSensor = false;
	return d;
}

m3d_hinge_joint_def m3d_hinge_joint_def_default(void)
{
	m3d_hinge_joint_def d;
	memset(&d, 0, sizeof(d));
	d.bodyA = m3d_body_id_null();
	d.bodyB = m3d_body_id_null();
	d.localAxisA = m3d_v3(0.0f, 1.0f, 0.0f);
	return d;
}

m3d_weld_joint_def m3d_weld_joint_def_default(void)
{
	m3d_weld_joint_def d;
	memset(&d, 0, sizeof(d));
	d.bodyA = m3d_body_id_null();
	d.bodyB = m3d_body_id_null();
	return d;
}

const char* m3dx_last_error(void) { return g_error; }

int m3dx_set_device(int device)
{
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
	{
		set_error("device %d not available (%d devices)", device, count);
		return -1;
	}
	g_device = device;
	return 0;
}

// The device radix sort on host arrays, on the device m3dx_set_device
// names, with buffers of its own (a test hook for the sort every stage uses).
int m3dx_sort_pairs(const uint64_t* keys, const uint32_t* vals, int64_t n, int bits, uint64_t* keysOut,
					uint32_t* valsOut)
{
	if (n < 0 || bits < 0 || bits > 64 || (n > 0 && (keys == nullptr || vals == nullptr || keysOut == nullptr ||
													valsOut == nullptr)))
	{
		set_error("sort pairs: bad arguments");
		return -1;
	}
	if (n == 0)
		return 0;
	if (!hip_ok(hipSetDevice(g_device), "sort pairs"))
		return -1;
	hipStream_t s = nullptr;
	uint64_t *k0 = nullptr, *k1 = nullptr, *status = nullptr;
	uint32_t *v0 = nullptr, *v1 = nullptr;
	void* scratch = nullptr;
	const size_t words = (size_t)std::max(scan_status_words(n), sort_status_words(n));
	bool ok = hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "sort pairs") &&
			  hip_ok(hipMalloc((void**)&k0, n * sizeof(uint64_t)), "sort pairs") &&
			  hip_ok(hipMalloc((void**)&k1, n * sizeof(uint64_t)), "sort pairs") &&
			  hip_ok(hipMalloc((void**)&v0, n * sizeof(uint32_t)), "sort pairs") &&
			  hip_ok(hipMalloc((void**)&v1, n * sizeof(uint32_t)), "sort pairs") &&
			  hip_ok(hipMalloc(&scratch, sort_scratch_bytes(n)), "sort pairs") &&
			  hip_ok(hipMalloc((void**)&status, words * sizeof(uint64_t)), "sort pairs") &&
			  hip_ok(hipMemset(status, 0, words * sizeof(uint64_t)), "sort pairs") &&
			  hip_ok(hipMemcpyAsync(k0, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice, s), "sort pairs") &&
			  hip_ok(hipMemcpyAsync(v0, vals, n * sizeof(uint32_t), hipMemcpyHostToDevice, s), "sort pairs");
	if (ok)
	{
		radix_sort_pairs(k0, v0, k1, v1, n, bits, scratch, status, s);
		ok = hip_ok(hipGetLastError(), "sort pairs") &&
			 hip_ok(hipMemcpyAsync(keysOut, k0, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "sort pairs") &&
			 hip_ok(hipMemcpyAsync(valsOut, v0, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "sort pairs") &&
			 hip_ok(hipStreamSynchronize(s), "sort pairs");
	}
	for (void* p : { (void*)k0, (void*)k1, (void*)v0, (void*)v1, scratch, (void*)status })
		if (p != nullptr)
			(void)hipFree(p);
	if (s != nullptr)
		(void)hipStreamDestroy(s);
	return ok ? 0 : -1;
}

// src/world.cpp:468-500; returns NULL when the device cannot be set up
m3d_world* m3d_world_create(const m3d_world_def* def)
{
	m3d_world* wp = nullptr;
	try
	{
		wp = new m3d_world();
	}
	catch (...)
	{
		return nullptr;
	}
	World& w = wp->w;
	w.def = def != nullptr ? *def : m3d_world_def_default();
	if (w.def.workerCount < 1)
		w.def.workerCount = 1;
	if (w.def.gridCellSize <= 0.1f)
		w.def.gridCellSize = 2.5f;
	w.device = g_device;
	bool ok = hip_ok(hipSetDevice(w.device), "hipSetDevice") &&
			  hip_ok(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking), "stream") &&
			  w.counters.ensure(kCountersSize) &&
			  hip_ok(hipHostMalloc((void**)&w.hostCounters, kHostCountersWords * sizeof(uint32_t), hipHostMallocDefault),
					 "pinned counters");
	if (!ok)
	{
		delete wp;
		return nullptr;
	}
	memset(w.hostCounters, 0, kHostCountersWords * sizeof(uint32_t));
	memset(&w.profile, 0, sizeof(w.profile));
	return wp;
}

// Batched worlds: one device world holding `count` independent sub-worlds
// that share the world definition. Bodies go to the selected sub-world;
// sub-worlds never interact and each steps exactly as a world of its own.
m3d_world* m3dx_batch_create(const m3d_world_def* def, int32_t count)
{
	if (count < 1 || count > (1 << 18))
	{
		set_error("batch size %d outside [1, 262144]", count);
		return nullptr;
	}
	m3d_world* wp = m3d_world_create(def);
	if (wp == nullptr)
		return nullptr;
	World& w = wp->w;
	w.subCount = count;
	w.subCurrent = 0;
	w.subBodyStart.assign((size_t)count + 1, 0);
	w.subDirtyHost.assign((size_t)count, 0);
	w.subStableHost.assign((size_t)count, 0);
	bool ok = w.subStable.ensure(count) && w.subContactStart.ensure((size_t)count + 1) &&
			  w.subWork.ensure(4 * (size_t)count + 4) && w.subDirtyUpload.ensure(count);
	if (ok)
	{
		// every sub-world starts dirty, 1000 steps after its last stable build
		// (src/m3d_internal.h:624-626)
		std::vector<uint32_t> init((size_t)count, SUB_DIRTY | (1000u << 8));
		ok = hip_ok(hipMemcpy(w.subStable.p, init.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice),
					"batch init") &&
			 hip_ok(hipMemset(w.subContactStart.p, 0, ((size_t)count + 1) * sizeof(int32_t)), "batch init");
	}
	if (!ok)
	{
		m3d_world_destroy(wp);
		return nullptr;
	}
	return wp;
}

// Selects the sub-world that m3d_body_create adds bodies to; sub-worlds are
// filled in ascending order. Returns 0, or -1 (with m3dx_last_error) for a
// plain world or an index below the current one.
int m3dx_batch_select(m3d_world* world, int32_t sub)
{
	World& w = world->w;
	if (w.subCount == 0 || sub < w.subCurrent || sub >= w.subCount)
	{
		set_error("sub-world %d cannot be selected (batch of %d, current %d)", sub, w.subCount, w.subCurrent);
		return -1;
	}
	w.subCurrent = sub;
	return 0;
}

int32_t m3dx_batch_size(m3d_world* world) { return world->w.subCount; }

// Body index range of a sub-world; returns -1 for a bad index.
int m3dx_batch_body_range(m3d_world* world, int32_t sub, int32_t* first, int32_t* count)
{
	World& w = world->w;
	if (sub < 0 || sub >= w.subCount)
		return -1;
	*first = w.subBodyStart[sub];
	*count = w.subBodyStart[sub + 1] - w.subBodyStart[sub];
	return 0;
}

// Contacts per sub-world after the last step (cap entries); returns the
// number of sub-worlds.
int m3dx_batch_contact_counts(m3d_world* world, int32_t* out, int32_t cap)
{
	World& w = world->w;
	if (w.subCount == 0)
		return 0;
	std::vector<int32_t> st((size_t)w.subCount + 1);
	(void)hipStreamSynchronize(w.stream);
	if (!hip_ok(hipMemcpy(st.data(), w.subContactStart.p, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost),
				"batch contact counts"))
		return -1;
	for (int32_t s = 0; s < w.subCount && s < cap; ++s)
		out[s] = st[s + 1] - st[s];
	return w.subCount;
}

void m3d_world_destroy(m3d_world* world)
{
	if (world == nullptr)
		return;
	(void)hipStreamSynchronize(world->w.stream);
	delete world;
}

void m3d_world_step(m3d_world* world, float dt, int substepCount)
{
	if (world == nullptr)
		return;
	world_step(world->w, dt, substepCount);
}

// src/world.cpp:512-567
m3d_body_id m3d_body_create(m3d_world* world, const m3d_body_def* bd, const m3d_shape_def* sd)
{
	World& w = world->w;
	int32_t index;
	// a batch reuses a destroyed slot of the sub-world being filled (each
	// sub-world's own free list, as its own world would)
	std::vector<int32_t>& freeList =
		w.subCount > 0 && w.subCurrent < (int32_t)w.subFreeBodies.size() ? w.subFreeBodies[w.subCurrent] : w.freeBodies;
	const bool reused = !freeList.empty();
	if (reused)
	{
		index = freeList.back();
		freeList.pop_back();
	}
	else
	{
		index = (int32_t)w.bodies.size();
		w.bodies.emplace_back();
		memset(&w.bodies[index], 0, sizeof(HostBody));
		w.bodies[index].generation = 1;
	}
	HostBody& b = w.bodies[index];
	uint32_t gen = b.generation;
	memset(&b, 0, sizeof(HostBody));
	b.rotation = m3d_quat_identity();
	b.prevRotation = m3d_quat_identity();
	b.generation = gen;
	b.inUse = true;
	b.type = bd->type;
	b.position = bd->position;
	b.rotation = m3d_quat_normalize(bd->orientation);
	b.linearVelocity = bd->linearVelocity;
	b.angularVelocity = bd->angularVelocity;
	b.linearDamping = bd->linearDamping;
	b.angularDamping = bd->angularDamping;
	b.gravityScale = bd->gravityScale;
	b.enableSleep = bd->enableSleep && w.def.enableSleep;
	b.isAwake = bd->type == M3D_BODY_STATIC ? false : bd->isAwake;
	b.sleepTime = 0.0f;
	b.smoothedEnergy = 4.0f * kSleepEnergyThresh;
	b.userData = bd->userData;
	b.shape.type = (int32_t)sd->type;
	b.shape.radius = sd->radius;
	b.shape.halfHeight = sd->halfHeight;
	b.shape.halfExtents = sd->halfExtents;
	b.friction = sd->friction;
	b.restitution = sd->restitution;
	b.rollingResistance = sd->rollingResistance;
	b.filter = sd->filter;
	b.isSensor = sd->isSensor;
	compute_mass(b, sd);
	update_inv_inertia(b);
	b.bp = 0; // recycled slot: stale fat AABB and tier membership reset
	w.gridDirty = true;
	if (w.subCount > 0 && !reused)
	{
		// bodies of a sub-world are one index range, in sub-world order
		w.bodySubHost.push_back(w.subCurrent);
		for (int32_t s = w.subCurrent + 1; s <= w.subCount; ++s)
			w.subBodyStart[s] = index + 1;
		w.bodySubDirty = true;
	}
	set_stable_dirty(w, index);
	mark_dirty(w, index);
	return make_id(w, index);
}

// src/world.cpp:569-615
// Wakes a destroyed body's partner: on the host record when that record is
// current (the mirror is fresh, or the body is dirty and its record will be
// uploaded), else through the device wake list (only the wake fields change).
static void wake_partner(World& w, int32_t i)
{
	if (!w.hostStale || (w.dirtyFlag.size() > (size_t)i && w.dirtyFlag[i]))
	{
		wake_body_host(w, i);
		return;
	}
	if (w.bodies[i].type == M3D_BODY_STATIC)
		return;
	set_stable_dirty(w, i);
	w.deviceWake.push_back(i);
}

static void retire_body(World& w, int32_t index);

void m3d_body_destroy(m3d_world* world, m3d_body_id id)
{
	World& w = world->w;
	HostBody* b = body_checked(w, id);
	if (b == nullptr)
		return;
	int32_t index = id.index;
	if (w.subCount > 0)
	{
		// the body's sub-world as a world of its own: its contacts leave the
		// sub-world's run on the device (touching ones give end events and,
		// unless sensors, wake the partner), then the slot joins the
		// sub-world's free list and its stable tier is rebuilt at the next step
		std::vector<int4> gone;
		if (!batch_destroy_contacts(w, index, gone))
			return;
		for (const int4& e : gone)
		{
			int32_t other = e.x == index ? e.y : e.x;
			if (e.z == 0)
				wake_partner(w, other);
			m3d_contact_event ev;
			ev.bodyA = make_id(w, e.x);
			ev.bodyB = make_id(w, e.y);
			ev.isSensor = e.z != 0;
			w.endEvents.push_back(ev);
		}
		retire_body(w, index);
		return;
	}
	w.destroyedSinceStep = true;
	if (ensure_sidecar(w))
	{
		// contacts of this body, removed in descending index order with
		// swap-with-last (src/world.cpp:576-593, :100-119). A descending scan
		// only ever swaps in a contact it has already passed, so the body's
		// contacts go in descending order of their current positions.
		std::vector<int32_t> mine;
		if (index < w.hcN)
		{
			for (int32_t s = w.hcStart[index]; s < w.hcStart[index + 1]; ++s)
				if (w.hcList[s] >= 0)
					mine.push_back(w.hcList[s]);
		}
		std::sort(mine.begin(), mine.end(), [](int32_t a, int32_t b) { return a > b; });
		for (int32_t i : mine)
		{
			int4 e = w.hcSide[i];
			int2 c = make_int2(e.x, e.y);
			int32_t other = c.x == index ? c.y : c.x;
			uint32_t cf = (uint32_t)e.z;
			if (cf & CF_TOUCHING)
			{
				if (!(cf & CF_SENSOR))
					wake_partner(w, other);
				m3d_contact_event ev;
				ev.bodyA = make_id(w, c.x);
				ev.bodyB = make_id(w, c.y);
				ev.isSensor = (cf & CF_SENSOR) != 0;
				w.endEvents.push_back(ev);
			}
			int2 gone = w.hcSlot[i];
			if (gone.x >= 0)
				w.hcList[gone.x] = -1;
			if (gone.y >= 0)
				w.hcList[gone.y] = -1;
			int32_t last = w.hcCount - 1;
			if (i != last)
			{
				w.hcSide[i] = w.hcSide[last];
				int2 sl = w.hcSlot[last];
				w.hcSlot[i] = sl;
				if (sl.x >= 0)
					w.hcList[sl.x] = i;
				if (sl.y >= 0)
					w.hcList[sl.y] = i;
				w.hcMoved.push_back(i);
			}
			--w.hcCount;
			w.pendingPermutation = true;
		}
	}
	retire_body(w, index);
}

// The end of m3d_body_destroy after its contacts (src/world.cpp:595-614):
// joints on the body go (their other body wakes), the slot is freed with its
// generation bumped, and the broadphase must rebuild both tiers.
static void retire_body(World& w, int32_t index)
{
	for (int32_t i = 0; i < (int32_t)w.joints.size(); ++i)
	{
		DJoint& j = w.joints[i];
		if (j.inUse && (j.bodyA == index || j.bodyB == index))
		{
			wake_body_host(w, j.bodyA == index ? j.bodyB : j.bodyA);
			j.inUse = 0;
			w.jointGen[i]++;
			w.freeJoints.push_back(i);
			w.jointedPairsDirty = true;
			w.jointsDirty = true;
		}
	}
	HostBody& b = w.bodies[index];
	b.inUse = false;
	b.generation++;
	w.gridDirty = true;
	if (w.subCount > 0)
	{
		// the sub-world's free list; its stable tier rebuilt at the next step
		int32_t sub = w.bodySubHost[index];
		if ((int32_t)w.subFreeBodies.size() < w.subCount)
			w.subFreeBodies.resize(w.subCount);
		w.subFreeBodies[sub].push_back(index);
		w.subDirtyHost[sub] |= (uint8_t)(SUB_DIRTY | SUB_FORCE);
		w.anySubDirtyHost = true;
	}
	else
	{
		w.freeBodies.push_back(index);
		w.stableDirty = true;
		w.stableForce = true;
	}
	mark_dirty(w, index);
}

// Bulk despawn and spawn: the calls m3d_body_destroy(destroy[k]) then
// ids[k] = m3d_body_create(&bodyDefs[k], &shapeDefs[k]) for k = 0 ..
// count - 1, in that order (src/world.cpp:512-615 per call). The slot
// bookkeeping (free list, generations, joints) runs in call order on the
// host; creates touch no contact, so the contact removals of all the
// destroys follow in one pass: the device lists the contacts touching the
// destroyed bodies, and the host replays the reference's removals on that
// list alone (per destroy, the body's contacts in descending current
// position, each swapped with the current last contact; only positions at
// or beyond the final count and the removed contacts below it take part),
// which yields the wakes, the end events in the reference's order, and the
// moves of the surviving tail contacts into the holes.
int m3dx_body_respawn(m3d_world* world, const m3d_body_id* destroy, const m3d_body_def* bodyDefs,
					  const m3d_shape_def* shapeDefs, int32_t count, m3d_body_id* ids)
{
	if (world == nullptr || count < 0 || (bodyDefs == nullptr) != (shapeDefs == nullptr) ||
		(bodyDefs != nullptr && ids == nullptr))
	{
		set_error("respawn: bad arguments");
		return -1;
	}
	World& w = world->w;
	if (destroy != nullptr && w.subCount > 0)
	{
		set_error("bodies of batched worlds cannot be destroyed");
		return -1;
	}
	if (count == 0)
		return 0;
	// the device contacts as the last step and the calls since left them
	if (destroy != nullptr && !push_host_changes(w))
		return -1;
	const int32_t nDev = w.deviceBodyCount;
	if ((int32_t)w.respRankHost.size() < (int32_t)w.bodies.size())
		w.respRankHost.resize(w.bodies.size(), -1);
	// (slot, rank) and the generation before the destroy of every body that
	// existed on the device when the call began
	std::vector<int2> destroyed;
	std::vector<uint32_t> genBefore;
	std::vector<int32_t> createdSlots;
	for (int32_t k = 0; k < count; ++k)
	{
		if (destroy != nullptr)
		{
			HostBody* b = body_checked(w, destroy[k]);
			if (b != nullptr)
			{
				int32_t index = destroy[k].index;
				// -2: created by this call (no contacts yet); >= 0: already destroyed
				if (index < nDev && w.respRankHost[index] == -1)
				{
					w.respRankHost[index] = (int32_t)destroyed.size();
					destroyed.push_back(make_int2(index, (int32_t)destroyed.size()));
					genBefore.push_back(b->generation);
				}
				retire_body(w, index);
			}
		}
		if (bodyDefs != nullptr)
		{
			ids[k] = m3d_body_create(world, &bodyDefs[k], &shapeDefs[k]);
			if ((int32_t)w.respRankHost.size() < (int32_t)w.bodies.size())
				w.respRankHost.resize(w.bodies.size(), -1);
			if (w.respRankHost[ids[k].index] == -1)
			{
				w.respRankHost[ids[k].index] = -2;
				createdSlots.push_back(ids[k].index);
			}
		}
	}
	auto reset = [&]() {
		for (const int2& d : destroyed)
			w.respRankHost[d.x] = -1;
		for (int32_t i : createdSlots)
			w.respRankHost[i] = -1;
	};
	if (destroyed.empty())
	{
		reset();
		return 0;
	}
	std::vector<int4> rem; // (body a, body b, flags, position), ascending position
	bool ok = gather_despawn_contacts(w, destroyed, rem);
	if (ok && !rem.empty())
	{
		const int32_t C = w.contactCount;
		const int32_t R = (int32_t)rem.size();
		const int32_t n1 = C - R; // final contact count
		auto rankOf = [&](int32_t b) { return w.respRankHost[b] >= 0 ? w.respRankHost[b] : INT32_MAX; };
		auto genOf = [&](int32_t b) {
			int32_t r = w.respRankHost[b];
			return r >= 0 ? genBefore[r] : w.bodies[b].generation;
		};
		// slots: tail slot t >= 0 is position n1 + t; hole h (encoded -1 - h)
		// is the position of removed entry h (entries below n1 come first)
		int32_t H = 0;
		while (H < R && rem[H].w < n1)
			++H;
		std::vector<int32_t> tailOcc((size_t)R), holeOcc((size_t)H), cur((size_t)R);
		for (int32_t t = 0; t < R; ++t)
			tailOcc[t] = R + t; // surviving tail contact, named by its tail slot
		for (int32_t i = 0; i < R; ++i)
		{
			if (i < H)
			{
				holeOcc[i] = i;
				cur[i] = -1 - i;
			}
			else
			{
				tailOcc[rem[i].w - n1] = i;
				cur[i] = rem[i].w - n1;
			}
		}
		auto posOf = [&](int32_t slot) { return slot >= 0 ? n1 + slot : rem[-1 - slot].w; };
		// entries grouped by the destroy that removes them (the earlier of
		// their two bodies' destroys)
		std::vector<int32_t> groupStart(destroyed.size() + 1, 0), order((size_t)R);
		for (int32_t i = 0; i < R; ++i)
			++groupStart[std::min(rankOf(rem[i].x), rankOf(rem[i].y)) + 1];
		for (size_t g = 0; g < destroyed.size(); ++g)
			groupStart[g + 1] += groupStart[g];
		{
			std::vector<int32_t> fill(groupStart.begin(), groupStart.end() - 1);
			for (int32_t i = 0; i < R; ++i)
				order[fill[std::min(rankOf(rem[i].x), rankOf(rem[i].y))]++] = i;
		}
		int32_t size = C;
		std::vector<std::pair<int32_t, int32_t>> grp;
		for (size_t g = 0; g < destroyed.size(); ++g)
		{
			const int32_t body = destroyed[g].x;
			grp.clear();
			for (int32_t q = groupStart[g]; q < groupStart[g + 1]; ++q)
				grp.emplace_back(posOf(cur[order[q]]), order[q]);
			std::sort(grp.begin(), grp.end(), [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
				return a.first > b.first;
			});
			for (const auto& pe : grp)
			{
				const int32_t i = pe.second;
				const int4 e = rem[i];
				const uint32_t cf = (uint32_t)e.z;
				if (cf & CF_TOUCHING)
				{
					const int32_t other = e.x == body ? e.y : e.x;
					// a partner destroyed later in this call is overwritten by
					// its own destroy (and a create in its slot)
					if (!(cf & CF_SENSOR) && w.respRankHost[other] < 0)
						wake_partner(w, other);
					m3d_contact_event ev;
					ev.bodyA = { e.x, genOf(e.x) };
					ev.bodyB = { e.y, genOf(e.y) };
					ev.isSensor = (cf & CF_SENSOR) != 0;
					w.endEvents.push_back(ev);
				}
				// swap with the current last contact
				const int32_t lastSlot = size - 1 - n1;
				const int32_t o = tailOcc[lastSlot];
				if (o != i)
				{
					const int32_t sl = cur[i];
					if (sl >= 0)
						tailOcc[sl] = o;
					else
						holeOcc[-1 - sl] = o;
					if (o < R)
						cur[o] = sl;
				}
				--size;
			}
		}
		std::vector<int2> mv;
		mv.reserve(H);
		for (int32_t h = 0; h < H; ++h)
		{
			const int32_t o = holeOcc[h];
			if (o < R)
			{
				set_error("respawn: contact removal replay left a hole unfilled");
				ok = false;
				break;
			}
			mv.push_back(make_int2(rem[h].w, n1 + (o - R)));
		}
		// (a copy from pageable memory has left the vector when it returns;
		// the moves run in stream order before the next use of the contacts)
		if (ok && !mv.empty())
			ok = w.moves.ensure(mv.size()) &&
				 hip_ok(hipMemcpyAsync(w.moves.p, mv.data(), mv.size() * sizeof(int2), hipMemcpyHostToDevice, w.stream),
						"respawn moves") &&
				 apply_moves(w, w.moves.p, (int32_t)mv.size());
		if (ok)
			w.contactCount = n1;
	}
	reset();
	w.sidecarValid = false;
	w.sidecarPrefetched = false;
	return ok ? 0 : -1;
}

bool m3d_body_is_valid(m3d_world* world, m3d_body_id id) { return body_checked(world->w, id) != nullptr; }

m3d_vec3 m3d_body_position(m3d_world* world, m3d_body_id id)
{
	ensure_host(world->w);
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->position : m3d_v3_zero();
}

m3d_quat m3d_body_rotation(m3d_world* world, m3d_body_id id)
{
	ensure_host(world->w);
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->rotation : m3d_quat_identity();
}

m3d_transform m3d_body_transform(m3d_world* world, m3d_body_id id)
{
	ensure_host(world->w);
	HostBody* b = body_checked(world->w, id);
	if (b == nullptr)
	{
		m3d_transform t = { m3d_v3_zero(), m3d_quat_identity() };
		return t;
	}
	m3d_transform t = { b->position, b->rotation };
	return t;
}

// src/world.cpp:644-661
void m3d_body_set_transform(m3d_world* world, m3d_body_id id, m3d_vec3 position, m3d_quat rotation)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr)
		return;
	b->position = position;
	b->rotation = m3d_quat_normalize(rotation);
	update_inv_inertia(*b);
	b->bp &= (uint8_t)~BP_HAS_FAT;
	w.gridDirty = true;
	set_stable_dirty(w, id.index);
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

m3d_vec3 m3d_body_linear_velocity(m3d_world* world, m3d_body_id id)
{
	ensure_host(world->w);
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->linearVelocity : m3d_v3_zero();
}

m3d_vec3 m3d_body_angular_velocity(m3d_world* world, m3d_body_id id)
{
	ensure_host(world->w);
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->angularVelocity : m3d_v3_zero();
}

void m3d_body_set_linear_velocity(m3d_world* world, m3d_body_id id, m3d_vec3 v)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr || b->type == M3D_BODY_STATIC)
		return;
	b->linearVelocity = v;
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

void m3d_body_set_angular_velocity(m3d_world* world, m3d_body_id id, m3d_vec3 av)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr || b->type == M3D_BODY_STATIC)
		return;
	b->angularVelocity = av;
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

// src/world.cpp:697-709
void m3d_body_apply_impulse(m3d_world* world, m3d_body_id id, m3d_vec3 impulse, m3d_vec3 worldPoint)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr || b->type != M3D_BODY_DYNAMIC)
		return;
	update_inv_inertia(*b);
	b->linearVelocity = m3d_mul_add(b->linearVelocity, b->invMass, impulse);
	m3d_vec3 r = m3d_sub(worldPoint, b->position);
	b->angularVelocity = m3d_add(b->angularVelocity, m3d_mat3_mulv(&b->invInertiaWorld, m3d_cross(r, impulse)));
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

float m3d_body_mass(m3d_world* world, m3d_body_id id)
{
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->mass : 0.0f;
}

bool m3d_body_is_awake(m3d_world* world, m3d_body_id id)
{
	ensure_host(world->w);
	HostBody* b = body_checked(world->w, id);
	return b != nullptr && b->isAwake;
}

void m3d_body_wake(m3d_world* world, m3d_body_id id)
{
	World& w = world->w;
	ensure_host(w);
	if (body_checked(w, id) != nullptr)
		wake_body_host(w, id.index);
}

uint64_t m3d_body_user_data(m3d_world* world, m3d_body_id id)
{
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->userData : 0;
}

void m3d_body_set_user_data(m3d_world* world, m3d_body_id id, uint64_t userData)
{
	HostBody* b = body_checked(world->w, id);
	if (b != nullptr)
		b->userData = userData; // host-only field
}

m3d_body_type m3d_body_get_type(m3d_world* world, m3d_body_id id)
{
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->type : M3D_BODY_STATIC;
}

// src/world.cpp:753-801
void m3d_body_apply_force(m3d_world* world, m3d_body_id id, m3d_vec3 force)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr || b->type != M3D_BODY_DYNAMIC)
		return;
	b->force = m3d_add(b->force, force);
	w.anyForcesApplied = true;
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

void m3d_body_apply_force_at_point(m3d_world* world, m3d_body_id id, m3d_vec3 force, m3d_vec3 worldPoint)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr || b->type != M3D_BODY_DYNAMIC)
		return;
	b->force = m3d_add(b->force, force);
	m3d_vec3 r = m3d_sub(worldPoint, b->position);
	b->torque = m3d_add(b->torque, m3d_cross(r, force));
	w.anyForcesApplied = true;
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

void m3d_body_apply_torque(m3d_world* world, m3d_body_id id, m3d_vec3 torque)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr || b->type != M3D_BODY_DYNAMIC)
		return;
	b->torque = m3d_add(b->torque, torque);
	w.anyForcesApplied = true;
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

void m3d_body_apply_angular_impulse(m3d_world* world, m3d_body_id id, m3d_vec3 impulse)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr || b->type != M3D_BODY_DYNAMIC)
		return;
	update_inv_inertia(*b);
	b->angularVelocity = m3d_add(b->angularVelocity, m3d_mat3_mulv(&b->invInertiaWorld, impulse));
	mark_dirty(w, id.index);
	wake_body_host(w, id.index);
}

float m3d_body_friction(m3d_world* world, m3d_body_id id)
{
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->friction : 0.0f;
}

static void refresh_materials(World& w, int32_t index)
{
	if ((int32_t)w.materialRefresh.size() < (int32_t)w.bodies.size())
		w.materialRefresh.resize(w.bodies.size(), 0);
	w.materialRefresh[index] = 1;
	w.anyMaterialRefresh = true;
}

// src/world.cpp:826-852
void m3d_body_set_friction(m3d_world* world, m3d_body_id id, float friction)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr)
		return;
	b->friction = friction < 0.0f ? 0.0f : friction;
	mark_dirty(w, id.index);
	refresh_materials(w, id.index);
}

float m3d_body_restitution(m3d_world* world, m3d_body_id id)
{
	HostBody* b = body_checked(world->w, id);
	return b != nullptr ? b->restitution : 0.0f;
}

void m3d_body_set_restitution(m3d_world* world, m3d_body_id id, float restitution)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* b = body_checked(w, id);
	if (b == nullptr)
		return;
	b->restitution = restitution < 0.0f ? 0.0f : restitution;
	mark_dirty(w, id.index);
	refresh_materials(w, id.index);
}

m3d_vec3 m3d_world_gravity(m3d_world* world) { return world->w.def.gravity; }

void m3d_world_set_gravity(m3d_world* world, m3d_vec3 gravity) { world->w.def.gravity = gravity; }

// ------------------------------------------------------------ joints

// src/world.cpp:868-890
// joints of batched worlds stay inside one sub-world
static bool same_sub_world(const World& w, int32_t a, int32_t b)
{
	return w.subCount == 0 || w.bodySubHost[a] == w.bodySubHost[b];
}

static m3d_joint_id allocate_joint(World& w, DJoint** out)
{
	int32_t index;
	if (!w.freeJoints.empty())
	{
		index = w.freeJoints.back();
		w.freeJoints.pop_back();
	}
	else
	{
		index = (int32_t)w.joints.size();
		w.joints.emplace_back();
		w.jointGen.push_back(1);
	}
	DJoint& j = w.joints[index];
	memset(&j, 0, sizeof(DJoint));
	j.relRot0 = m3d_quat_identity();
	j.inUse = 1;
	w.jointedPairsDirty = true;
	w.jointsDirty = true;
	*out = &j;
	return { index, w.jointGen[index] };
}

m3d_joint_id m3d_joint_create_distance(m3d_world* world, const m3d_distance_joint_def* def)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* a = body_checked(w, def->bodyA);
	HostBody* b = body_checked(w, def->bodyB);
	if (a == nullptr || b == nullptr || def->bodyA.index == def->bodyB.index ||
		!same_sub_world(w, def->bodyA.index, def->bodyB.index))
		return { -1, 0 };
	DJoint* j;
	m3d_joint_id id = allocate_joint(w, &j);
	j->type = 0;
	j->bodyA = def->bodyA.index;
	j->bodyB = def->bodyB.index;
	j->localAnchorA = def->localAnchorA;
	j->localAnchorB = def->localAnchorB;
	j->hertz = def->hertz;
	j->dampingRatio = def->dampingRatio;
	j->lambda = 0.0f;
	if (def->length > 0.0f)
	{
		j->length = def->length;
	}
	else
	{
		m3d_transform ta = { a->position, a->rotation }, tb = { b->position, b->rotation };
		m3d_vec3 pA = m3d_transform_point(ta, def->localAnchorA);
		m3d_vec3 pB = m3d_transform_point(tb, def->localAnchorB);
		j->length = m3d_length(m3d_sub(pB, pA));
	}
	wake_body_host(w, j->bodyA);
	wake_body_host(w, j->bodyB);
	return id;
}

m3d_joint_id m3d_joint_create_ball(m3d_world* world, const m3d_ball_joint_def* def)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* a = body_checked(w, def->bodyA);
	HostBody* b = body_checked(w, def->bodyB);
	if (a == nullptr || b == nullptr || def->bodyA.index == def->bodyB.index ||
		!same_sub_world(w, def->bodyA.index, def->bodyB.index))
		return { -1, 0 };
	DJoint* j;
	m3d_joint_id id = allocate_joint(w, &j);
	j->type = 1;
	j->bodyA = def->bodyA.index;
	j->bodyB = def->bodyB.index;
	j->localAnchorA = def->localAnchorA;
	j->localAnchorB = def->localAnchorB;
	j->lambda = 0.0f;
	wake_body_host(w, j->bodyA);
	wake_body_host(w, j->bodyB);
	return id;
}

// src/world.cpp:946-996
m3d_joint_id m3d_joint_create_hinge(m3d_world* world, const m3d_hinge_joint_def* def)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* a = body_checked(w, def->bodyA);
	HostBody* b = body_checked(w, def->bodyB);
	if (a == nullptr || b == nullptr || def->bodyA.index == def->bodyB.index ||
		!same_sub_world(w, def->bodyA.index, def->bodyB.index))
		return { -1, 0 };
	DJoint* j;
	m3d_joint_id id = allocate_joint(w, &j);
	j->type = 2;
	j->bodyA = def->bodyA.index;
	j->bodyB = def->bodyB.index;
	j->localAnchorA = def->localAnchorA;
	j->localAnchorB = def->localAnchorB;
	j->localAxisA = m3d_normalize(def->localAxisA);
	if (m3d_length_sq(j->localAxisA) < 0.5f)
		j->localAxisA = m3d_v3(0.0f, 1.0f, 0.0f);
	if (m3d_length_sq(def->localAxisB) > 0.5f)
	{
		j->localAxisB = m3d_normalize(def->localAxisB);
	}
	else
	{
		m3d_vec3 axisWorld = m3d_rotate(a->rotation, j->localAxisA);
		j->localAxisB = m3d_inv_rotate(b->rotation, axisWorld);
	}
	{
		m3d_vec3 axisWorld = m3d_rotate(a->rotation, j->localAxisA);
		m3d_vec3 pick = fabsf(axisWorld.x) < 0.9f ? m3d_v3(1.0f, 0.0f, 0.0f) : m3d_v3(0.0f, 1.0f, 0.0f);
		m3d_vec3 refWorld = m3d_normalize(m3d_cross(axisWorld, pick));
		j->localRefA = m3d_inv_rotate(a->rotation, refWorld);
		j->localRefB = m3d_inv_rotate(b->rotation, refWorld);
	}
	j->enableLimit = def->enableLimit;
	j->lowerAngle = def->lowerAngle;
	j->upperAngle = def->upperAngle;
	j->enableMotor = def->enableMotor;
	j->motorSpeed = def->motorSpeed;
	j->maxMotorTorque = def->maxMotorTorque;
	j->lambda = 0.0f;
	j->motorLambda = 0.0f;
	wake_body_host(w, j->bodyA);
	wake_body_host(w, j->bodyB);
	return id;
}

m3d_joint_id m3d_joint_create_weld(m3d_world* world, const m3d_weld_joint_def* def)
{
	World& w = world->w;
	ensure_host(w);
	HostBody* a = body_checked(w, def->bodyA);
	HostBody* b = body_checked(w, def->bodyB);
	if (a == nullptr || b == nullptr || def->bodyA.index == def->bodyB.index ||
		!same_sub_world(w, def->bodyA.index, def->bodyB.index))
		return { -1, 0 };
	DJoint* j;
	m3d_joint_id id = allocate_joint(w, &j);
	j->type = 3;
	j->bodyA = def->bodyA.index;
	j->bodyB = def->bodyB.index;
	j->localAnchorA = def->localAnchorA;
	j->localAnchorB = def->localAnchorB;
	j->relRot0 = m3d_quat_normalize(m3d_quat_mul(m3d_quat_conj(a->rotation), b->rotation));
	j->lambda = 0.0f;
	wake_body_host(w, j->bodyA);
	wake_body_host(w, j->bodyB);
	return id;
}

float m3d_joint_hinge_angle(m3d_world* world, m3d_joint_id id)
{
	World& w = world->w;
	ensure_host(w);
	DJoint* j = joint_checked(w, id);
	if (j == nullptr || j->type != 2)
		return 0.0f;
	return hinge_angle(w.bodies[j->bodyA].rotation, w.bodies[j->bodyB].rotation, *j);
}

void m3d_joint_hinge_set_motor(m3d_world* world, m3d_joint_id id, bool enable, float motorSpeed, float maxMotorTorque)
{
	World& w = world->w;
	ensure_host(w);
	DJoint* j = joint_checked(w, id);
	if (j == nullptr || j->type != 2)
		return;
	j->enableMotor = enable;
	j->motorSpeed = motorSpeed;
	j->maxMotorTorque = maxMotorTorque < 0.0f ? 0.0f : maxMotorTorque;
	w.jointsDirty = true;
	wake_body_host(w, j->bodyA);
	wake_body_host(w, j->bodyB);
}

void m3d_joint_destroy(m3d_world* world, m3d_joint_id id)
{
	World& w = world->w;
	if (id.index < 0 || id.index >= (int32_t)w.joints.size())
		return;
	DJoint& j = w.joints[id.index];
	if (!j.inUse || w.jointGen[id.index] != id.generation)
		return;
	ensure_host(w);
	wake_body_host(w, j.bodyA);
	wake_body_host(w, j.bodyB);
	j.inUse = 0;
	w.jointGen[id.index]++;
	w.freeJoints.push_back(id.index);
	w.jointedPairsDirty = true;
	w.jointsDirty = true;
}

// ------------------------------------------------------------ queries

static Shape shape_of(const HostBody& b) { return b.shape; }

// Closest hit over every in-use body (the reference narrows candidates with a
// DDA walk over the grid tiers, src/world.cpp:1081-1200, and takes the
// smallest fraction among them; the set of bodies a segment can hit is the
// same). A dirty broadphase is refreshed with dt = 0 first, as the
// reference does, because that refresh is observable state.
static m3d_ray_result ray_cast_internal(World& w, m3d_vec3 origin, m3d_vec3 translation, uint32_t maskBits)
{
	m3d_ray_result r;
	r.body = m3d_body_id_null();
	r.point = m3d_add(origin, translation);
	r.normal = m3d_v3_zero();
	r.fraction = 2.0f;
	r.hit = false;
	// (batched worlds: the refresh state is per sub-world, so queries read
	// the broadphase of the last step)
	if (w.subCount == 0 && (w.gridDirty || w.stableDirty))
	{
		if (push_host_changes(w))
		{
			(void)refresh_broadphase(w, 0.0f);
			(void)hipStreamSynchronize(w.stream);
			w.hostStale = true;
		}
	}
	ensure_host(w);
	float maxFraction = 1.0f;
	for (int32_t i = 0; i < (int32_t)w.bodies.size(); ++i)
	{
		const HostBody& b = w.bodies[i];
		if (!b.inUse || (b.filter.categoryBits & maskBits) == 0)
			continue;
		float f;
		m3d_vec3 n;
		m3d_transform xf = { b.position, b.rotation };
		if (ray_cast_shape(shape_of(b), xf, origin, translation, &f, &n) && f < maxFraction)
		{
			maxFraction = f;
			r.hit = true;
			r.body = make_id(w, i);
			r.fraction = f;
			r.point = m3d_mul_add(origin, f, translation);
			r.normal = n;
		}
	}
	return r;
}

m3d_ray_result m3d_world_ray_cast(m3d_world* world, m3d_vec3 origin, m3d_vec3 translation)
{
	return ray_cast_internal(world->w, origin, translation, 0xFFFFFFFFu);
}

m3d_ray_result m3d_world_ray_cast_filtered(m3d_world* world, m3d_vec3 origin, m3d_vec3 translation, uint32_t maskBits)
{
	return ray_cast_internal(world->w, origin, translation, maskBits);
}

// src/world.cpp:1286-1392
m3d_ray_result m3d_world_sphere_cast(m3d_world* world, m3d_vec3 origin, float radius, m3d_vec3 translation,
									 uint32_t maskBits)
{
	World& w = world->w;
	ensure_host(w);
	m3d_ray_result r;
	r.body = m3d_body_id_null();
	r.point = m3d_add(origin, translation);
	r.normal = m3d_v3_zero();
	r.fraction = 2.0f;
	r.hit = false;
	radius = radius > 0.0f ? radius : 0.0f;
	m3d_vec3 end = m3d_add(origin, translation);
	m3d_vec3 pad = m3d_v3(radius + 1e-3f, radius + 1e-3f, radius + 1e-3f);
	m3d_aabb sweep;
	sweep.lowerBound = m3d_sub(m3d_min(origin, end), pad);
	sweep.upperBound = m3d_add(m3d_max(origin, end), pad);
	float len = m3d_length(translation);
	for (int32_t bi = 0; bi < (int32_t)w.bodies.size(); ++bi)
	{
		const HostBody& b = w.bodies[bi];
		if (!b.inUse || (b.filter.categoryBits & maskBits) == 0)
			continue;
		m3d_transform xf = { b.position, b.rotation };
		if (!m3d_aabb_overlap(sweep, shape_aabb(b.shape, xf)))
			continue;
		{
			m3d_vec3 closest;
			float raw = point_shape_closest(b.shape, xf, origin, &closest);
			if (raw - radius < 1e-4f)
			{
				r.hit = true;
				r.body = make_id(w, bi);
				r.fraction = 0.0f;
				r.point = closest;
				m3d_vec3 outward = raw >= 0.0f ? m3d_sub(origin, closest) : m3d_sub(closest, origin);
				r.normal = m3d_normalize(outward);
				return r;
			}
		}
		if (b.shape.type == M3D_SHAPE_SPHERE || b.shape.type == M3D_SHAPE_CAPSULE)
		{
			Shape inflated = b.shape;
			inflated.radius += radius;
			float f;
			m3d_vec3 n;
			if (ray_cast_shape(inflated, xf, origin, translation, &f, &n) && (!r.hit || f < r.fraction))
			{
				m3d_vec3 center = m3d_mul_add(origin, f, translation);
				r.hit = true;
				r.body = make_id(w, bi);
				r.fraction = f;
				r.normal = n;
				r.point = m3d_mul_add(center, -radius, n);
			}
			continue;
		}
		float t = 0.0f;
		for (int it = 0; it < 128; ++it)
		{
			m3d_vec3 p = m3d_mul_add(origin, t, translation);
			m3d_vec3 closest;
			float raw = point_shape_closest(b.shape, xf, p, &closest);
			float d = raw - radius;
			if (d < 1e-4f)
			{
				if (!r.hit || t < r.fraction)
				{
					r.hit = true;
					r.body = make_id(w, bi);
					r.fraction = t;
					r.point = closest;
					m3d_vec3 outward = raw >= 0.0f ? m3d_sub(p, closest) : m3d_sub(closest, p);
					r.normal = m3d_normalize(outward);
				}
				break;
			}
			if (len < 1e-9f)
				break;
			t += d / len;
			if (t > 1.0f || (r.hit && t >= r.fraction))
				break;
		}
	}
	return r;
}

// src/world.cpp:1394-1416
void m3d_world_overlap_aabb(m3d_world* world, m3d_aabb aabb, m3d_overlap_fn fn, void* context)
{
	if (fn == NULL)
		return;
	World& w = world->w;
	// the matching bodies found on the device, then handed to the callback
	// in ascending index as the reference's scan does; a body the callback
	// destroyed before its turn is skipped, as the scan's in-use test would
	std::vector<uint64_t> pairs;
	if (!overlap_query(w, &aabb, 1, pairs))
		return;
	for (uint64_t p : pairs)
	{
		int32_t bi = (int32_t)(uint32_t)p;
		if (bi >= (int32_t)w.bodies.size() || !w.bodies[bi].inUse)
			continue;
		if (!fn(make_id(w, bi), context))
			return;
	}
}

m3d_contact_events m3d_world_contact_events(m3d_world* world)
{
	World& w = world->w;
	m3d_contact_events ev;
	ev.beginEvents = w.beginEvents.data();
	ev.beginCount = (int32_t)w.beginEvents.size();
	ev.endEvents = w.endEvents.data();
	ev.endCount = (int32_t)w.endEvents.size();
	return ev;
}

m3d_profile m3d_world_profile(m3d_world* world) { return world->w.profile; }

// ------------------------------------------------------------ snapshots

static void put3(float* d, m3d_vec3 v)
{
	d[0] = v.x;
	d[1] = v.y;
	d[2] = v.z;
}
static void put4(float* d, m3d_quat q)
{
	d[0] = q.x;
	d[1] = q.y;
	d[2] = q.z;
	d[3] = q.w;
}
static m3d_vec3 get3(const float* s) { return m3d_v3(s[0], s[1], s[2]); }
static m3d_quat get4(const float* s)
{
	m3d_quat q = { s[0], s[1], s[2], s[3] };
	return q;
}

static void joint_to_state(const DJoint& j, uint32_t gen, m3dx_joint_state& o)
{
	memset(&o, 0, sizeof(o));
	o.type = j.type;
	o.bodyA = j.bodyA;
	o.bodyB = j.bodyB;
	put3(o.localAnchorA, j.localAnchorA);
	put3(o.localAnchorB, j.localAnchorB);
	o.length = j.length;
	o.hertz = j.hertz;
	o.dampingRatio = j.dampingRatio;
	put3(o.localAxisA, j.localAxisA);
	put3(o.localAxisB, j.localAxisB);
	put3(o.localRefA, j.localRefA);
	put3(o.localRefB, j.localRefB);
	o.enableLimit = j.enableLimit;
	o.lowerAngle = j.lowerAngle;
	o.upperAngle = j.upperAngle;
	o.enableMotor = j.enableMotor;
	o.motorSpeed = j.motorSpeed;
	o.maxMotorTorque = j.maxMotorTorque;
	put4(o.relRot0, j.relRot0);
	o.compliance = j.compliance;
	o.lambda = j.lambda;
	o.motorLambda = j.motorLambda;
	o.inUse = j.inUse;
	o.generation = gen;
}

static void state_to_joint(const m3dx_joint_state& o, DJoint& j)
{
	memset(&j, 0, sizeof(j));
	j.type = o.type;
	j.bodyA = o.bodyA;
	j.bodyB = o.bodyB;
	j.inUse = o.inUse;
	j.localAnchorA = get3(o.localAnchorA);
	j.localAnchorB = get3(o.localAnchorB);
	j.length = o.length;
	j.hertz = o.hertz;
	j.dampingRatio = o.dampingRatio;
	j.localAxisA = get3(o.localAxisA);
	j.localAxisB = get3(o.localAxisB);
	j.localRefA = get3(o.localRefA);
	j.localRefB = get3(o.localRefB);
	j.enableLimit = o.enableLimit;
	j.lowerAngle = o.lowerAngle;
	j.upperAngle = o.upperAngle;
	j.enableMotor = o.enableMotor;
	j.motorSpeed = o.motorSpeed;
	j.maxMotorTorque = o.maxMotorTorque;
	j.relRot0 = get4(o.relRot0);
	j.compliance = o.compliance;
	j.lambda = o.lambda;
	j.motorLambda = o.motorLambda;
}

int m3dx_world_export(m3d_world* world, m3dx_snapshot* s)
{
	if (world == nullptr || s == nullptr)
		return -1;
	World& w = world->w;
	if (!push_host_changes(w))
		return -1;
	ensure_host(w);
	m3dx_world_header& h = s->header;
	memset(&h, 0, sizeof(h));
	int32_t n = (int32_t)w.bodies.size();
	int32_t C = w.contactCount;
	h.bodyCapacity = n;
	h.contactCount = C;
	h.jointCapacity = (int32_t)w.joints.size();
	h.freeBodyCount = (int32_t)w.freeBodies.size();
	h.freeJointCount = (int32_t)w.freeJoints.size();
	h.stableEntryCount = (int32_t)w.stableCount;
	h.stableOversizedCount = (int32_t)w.stableOversized.size();
	h.gridDirty = w.gridDirty;
	h.stableDirty = w.stableDirty;
	h.stableForce = w.stableForce;
	h.stepsSinceStableBuild = w.stepsSinceStableBuild;
	h.anyForcesApplied = w.anyForcesApplied;
	h.jointedPairsDirty = w.jointedPairsDirty;
	put3(h.gravity, w.def.gravity);
	h.workerCount = w.def.workerCount;
	h.positionIterations = w.def.positionIterations;
	h.gridCellSize = w.def.gridCellSize;
	h.enableSleep = w.def.enableSleep;
	if (s->bodies != nullptr)
	{
		for (int32_t i = 0; i < n; ++i)
		{
			const HostBody& b = w.bodies[i];
			m3dx_body_state& o = s->bodies[i];
			memset(&o, 0, sizeof(o));
			put3(o.position, b.position);
			put4(o.rotation, b.rotation);
			put3(o.linearVelocity, b.linearVelocity);
			put3(o.angularVelocity, b.angularVelocity);
			put3(o.prevPosition, b.prevPosition);
			put4(o.prevRotation, b.prevRotation);
			put3(o.prevLinearVelocity, b.prevLinearVelocity);
			put3(o.prevAngularVelocity, b.prevAngularVelocity);
			o.invMass = b.invMass;
			put3(o.invInertiaLocal, b.invInertiaLocal);
			put3(o.invInertiaWorld + 0, b.invInertiaWorld.cx);
			put3(o.invInertiaWorld + 3, b.invInertiaWorld.cy);
			put3(o.invInertiaWorld + 6, b.invInertiaWorld.cz);
			o.shapeType = b.shape.type;
			o.radius = b.shape.radius;
			o.halfHeight = b.shape.halfHeight;
			put3(o.halfExtents, b.shape.halfExtents);
			o.friction = b.friction;
			o.restitution = b.restitution;
			o.rollingResistance = b.rollingResistance;
			o.mass = b.mass;
			put3(o.force, b.force);
			put3(o.torque, b.torque);
			o.categoryBits = b.filter.categoryBits;
			o.maskBits = b.filter.maskBits;
			o.groupIndex = b.filter.groupIndex;
			o.isSensor = b.isSensor;
			o.type = (int32_t)b.type;
			o.linearDamping = b.linearDamping;
			o.angularDamping = b.angularDamping;
			o.gravityScale = b.gravityScale;
			o.sleepTime = b.sleepTime;
			o.smoothedEnergy = b.smoothedEnergy;
			o.enableSleep = b.enableSleep;
			o.isAwake = b.isAwake;
			o.inUse = b.inUse;
			o.generation = b.generation;
			o.userDataLo = (uint32_t)b.userData;
			o.userDataHi = (uint32_t)(b.userData >> 32);
			put3(o.fatAABB, b.fat.lowerBound);
			put3(o.fatAABB + 3, b.fat.upperBound);
			o.hasFatAABB = (b.bp & BP_HAS_FAT) ? 1 : 0;
			o.inStableTier = (b.bp & BP_IN_STABLE_TIER) ? 1 : 0;
			o.inStableCells = (b.bp & BP_IN_STABLE_CELLS) ? 1 : 0;
		}
	}
	if (s->contacts != nullptr && C > 0)
	{
		int32_t cap = w.contactCap;
		std::vector<int2> bodies(C);
		std::vector<float4> mat(C), nrm(C), cpa(C), cpb(C), cqa(C), cqb(C), cch(4 * (size_t)C);
		std::vector<float4> pts[9];
		DBuf<float4>* pcols[9] = { &w.cc.pA, &w.cc.pB, &w.cc.pJA, &w.cc.pJB, &w.cc.pRA, &w.cc.pRB, &w.cc.pFA,
								   &w.cc.pFB, &w.cc.pSlip };
		bool ok = hip_ok(hipMemcpy(bodies.data(), w.cc.bodies.p, C * sizeof(int2), hipMemcpyDeviceToHost), "export") &&
				  hip_ok(hipMemcpy(mat.data(), w.cc.mat.p, C * sizeof(float4), hipMemcpyDeviceToHost), "export") &&
				  hip_ok(hipMemcpy(nrm.data(), w.cc.normal.p, C * sizeof(float4), hipMemcpyDeviceToHost), "export") &&
				  hip_ok(hipMemcpy(cch.data(), w.cc.cache.p, 4 * (size_t)C * sizeof(float4), hipMemcpyDeviceToHost),
						 "export");
		for (int32_t c = 0; c < C && ok; ++c)
		{
			cpa[c] = cch[4 * (size_t)c];
			cqa[c] = cch[4 * (size_t)c + 1];
			cpb[c] = cch[4 * (size_t)c + 2];
			cqb[c] = cch[4 * (size_t)c + 3];
		}
		for (int k = 0; k < 9 && ok; ++k)
		{
			pts[k].resize((size_t)4 * cap);
			ok = hip_ok(hipMemcpy(pts[k].data(), pcols[k]->p, (size_t)4 * cap * sizeof(float4), hipMemcpyDeviceToHost),
						"export");
		}
		if (!ok)
			return -1;
		for (int32_t c = 0; c < C; ++c)
		{
			m3dx_contact_state& o = s->contacts[c];
			memset(&o, 0, sizeof(o));
			uint32_t cf = f2u(mat[c].w);
			o.bodyA = bodies[c].x;
			o.bodyB = bodies[c].y;
			o.friction = mat[c].x;
			o.restitution = mat[c].y;
			o.rollingResistance = mat[c].z;
			o.touching = (cf & CF_TOUCHING) != 0;
			o.isSensor = (cf & CF_SENSOR) != 0;
			o.tunnelGuard = (cf & CF_TUNNEL_GUARD) != 0;
			o.curved = (cf & CF_CURVED) != 0;
			o.hasCache = (cf & CF_HAS_CACHE) != 0;
			put3(o.cachePosA, f4_v3(cpa[c]));
			put3(o.cachePosB, f4_v3(cpb[c]));
			put4(o.cacheRotA, f4_q(cqa[c]));
			put4(o.cacheRotB, f4_q(cqb[c]));
			put3(o.normal, f4_v3(nrm[c]));
			o.pointCount = f2i(nrm[c].w);
			for (int j = 0; j < 4; ++j)
			{
				size_t p = (size_t)c * 4 + j;
				m3dx_point_state& op = o.points[j];
				float4 a = pts[0][p], b = pts[1][p], ja = pts[2][p], jb = pts[3][p], ra = pts[4][p], rb = pts[5][p];
				put3(op.localAnchorA, f4_v3(a));
				op.separationOffset = a.w;
				put3(op.localAnchorB, f4_v3(b));
				op.wN = b.w;
				put3(op.angJA, f4_v3(ja));
				op.lambdaN = ja.w;
				put3(op.angJB, f4_v3(jb));
				op.lambdaT = jb.w;
				put3(op.rA, f4_v3(ra));
				op.separation = ra.w;
				put3(op.rB, f4_v3(rb));
				op.id = f2u(rb.w);
				put3(op.frictionAnchorA, f4_v3(pts[6][p]));
				put3(op.frictionAnchorB, f4_v3(pts[7][p]));
				put3(op.slipDir, f4_v3(pts[8][p]));
			}
		}
	}
	if (s->joints != nullptr)
		for (int32_t i = 0; i < h.jointCapacity; ++i)
			joint_to_state(w.joints[i], w.jointGen[i], s->joints[i]);
	if (s->freeBodies != nullptr)
		for (int32_t i = 0; i < h.freeBodyCount; ++i)
			s->freeBodies[i] = w.freeBodies[i];
	if (s->freeJoints != nullptr)
		for (int32_t i = 0; i < h.freeJointCount; ++i)
			s->freeJoints[i] = w.freeJoints[i];
	if (s->stableEntries != nullptr && w.stableCount > 0)
	{
		std::vector<uint64_t> k(w.stableCount);
		std::vector<uint32_t> b(w.stableCount);
		if (!hip_ok(hipMemcpy(k.data(), w.stableKeys.p, k.size() * 8, hipMemcpyDeviceToHost), "export") ||
			!hip_ok(hipMemcpy(b.data(), w.stableBodies.p, b.size() * 4, hipMemcpyDeviceToHost), "export"))
			return -1;
		for (size_t i = 0; i < k.size(); ++i)
		{
			s->stableEntries[i].key = k[i];
			s->stableEntries[i].body = (int32_t)b[i];
			s->stableEntries[i].pad = 0;
		}
	}
	if (s->stableOversized != nullptr)
		for (int32_t i = 0; i < h.stableOversizedCount; ++i)
			s->stableOversized[i] = w.stableOversized[i];
	return 0;
}

m3d_world* m3dx_world_import(const m3dx_snapshot* s)
{
	if (s == nullptr)
		return nullptr;
	const m3dx_world_header& h = s->header;
	m3d_world_def def = m3d_world_def_default();
	def.gravity = get3(h.gravity);
	def.workerCount = h.workerCount;
	def.positionIterations = h.positionIterations;
	def.gridCellSize = h.gridCellSize;
	def.enableSleep = h.enableSleep != 0;
	m3d_world* world = m3d_world_create(&def);
	if (world == nullptr)
		return nullptr;
	World& w = world->w;
	int32_t n = h.bodyCapacity;
	w.bodies.resize(n);
	for (int32_t i = 0; i < n; ++i)
	{
		const m3dx_body_state& o = s->bodies[i];
		HostBody& b = w.bodies[i];
		memset(&b, 0, sizeof(b));
		b.position = get3(o.position);
		b.rotation = get4(o.rotation);
		b.linearVelocity = get3(o.linearVelocity);
		b.angularVelocity = get3(o.angularVelocity);
		b.prevPosition = get3(o.prevPosition);
		b.prevRotation = get4(o.prevRotation);
		b.prevLinearVelocity = get3(o.prevLinearVelocity);
		b.prevAngularVelocity = get3(o.prevAngularVelocity);
		b.invMass = o.invMass;
		b.invInertiaLocal = get3(o.invInertiaLocal);
		b.invInertiaWorld.cx = get3(o.invInertiaWorld + 0);
		b.invInertiaWorld.cy = get3(o.invInertiaWorld + 3);
		b.invInertiaWorld.cz = get3(o.invInertiaWorld + 6);
		b.shape.type = o.shapeType;
		b.shape.radius = o.radius;
		b.shape.halfHeight = o.halfHeight;
		b.shape.halfExtents = get3(o.halfExtents);
		b.friction = o.friction;
		b.restitution = o.restitution;
		b.rollingResistance = o.rollingResistance;
		b.mass = o.mass;
		b.force = get3(o.force);
		b.torque = get3(o.torque);
		b.filter.categoryBits = o.categoryBits;
		b.filter.maskBits = o.maskBits;
		b.filter.groupIndex = o.groupIndex;
		b.isSensor = o.isSensor != 0;
		b.type = (m3d_body_type)o.type;
		b.linearDamping = o.linearDamping;
		b.angularDamping = o.angularDamping;
		b.gravityScale = o.gravityScale;
		b.sleepTime = o.sleepTime;
		b.smoothedEnergy = o.smoothedEnergy;
		b.enableSleep = o.enableSleep != 0;
		b.isAwake = o.isAwake != 0;
		b.inUse = o.inUse != 0;
		b.generation = o.generation;
		b.userData = (uint64_t)o.userDataLo | ((uint64_t)o.userDataHi << 32);
		b.fat.lowerBound = get3(o.fatAABB);
		b.fat.upperBound = get3(o.fatAABB + 3);
		b.bp = (uint8_t)((o.hasFatAABB ? BP_HAS_FAT : 0) | (o.inStableTier ? BP_IN_STABLE_TIER : 0) |
						 (o.inStableCells ? BP_IN_STABLE_CELLS : 0));
		mark_dirty(w, i);
	}
	w.freeBodies.assign(s->freeBodies, s->freeBodies + h.freeBodyCount);
	w.joints.resize(h.jointCapacity);
	w.jointGen.resize(h.jointCapacity);
	for (int32_t i = 0; i < h.jointCapacity; ++i)
	{
		state_to_joint(s->joints[i], w.joints[i]);
		w.jointGen[i] = s->joints[i].generation;
	}
	w.freeJoints.assign(s->freeJoints, s->freeJoints + h.freeJointCount);
	w.gridDirty = h.gridDirty != 0;
	w.stableDirty = h.stableDirty != 0;
	w.stableForce = h.stableForce != 0;
	w.stepsSinceStableBuild = h.stepsSinceStableBuild;
	w.anyForcesApplied = h.anyForcesApplied != 0;
	w.jointedPairsDirty = true;
	w.jointsDirty = true;
	if (!upload_dirty(w))
	{
		m3d_world_destroy(world);
		return nullptr;
	}
	// contacts
	int32_t C = h.contactCount;
	if (C > 0)
	{
		if (!ensure_device_contacts(w, C))
		{
			m3d_world_destroy(world);
			return nullptr;
		}
		int32_t cap = w.contactCap;
		std::vector<int2> bodies(C);
		std::vector<float4> mat(C), nrm(C), cch(4 * (size_t)C);
		std::vector<float4> pts[9];
		for (int k = 0; k < 9; ++k)
			pts[k].assign((size_t)4 * cap, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
		for (int32_t c = 0; c < C; ++c)
		{
			const m3dx_contact_state& o = s->contacts[c];
			uint32_t cf = (o.touching ? CF_TOUCHING : 0) | (o.isSensor ? CF_SENSOR : 0) |
						  (o.tunnelGuard ? CF_TUNNEL_GUARD : 0) | (o.curved ? CF_CURVED : 0) |
						  (o.hasCache ? CF_HAS_CACHE : 0);
			bodies[c] = make_int2(o.bodyA, o.bodyB);
			mat[c] = make_float4(o.friction, o.restitution, o.rollingResistance, u2f(cf));
			nrm[c] = v3_f4(get3(o.normal), i2f(o.pointCount));
			cch[4 * (size_t)c] = v3_f4(get3(o.cachePosA), 0.0f);
			cch[4 * (size_t)c + 1] = q_f4(get4(o.cacheRotA));
			cch[4 * (size_t)c + 2] = v3_f4(get3(o.cachePosB), 0.0f);
			cch[4 * (size_t)c + 3] = q_f4(get4(o.cacheRotB));
			for (int j = 0; j < 4; ++j)
			{
				size_t p = (size_t)c * 4 + j;
				const m3dx_point_state& op = o.points[j];
				pts[0][p] = v3_f4(get3(op.localAnchorA), op.separationOffset);
				pts[1][p] = v3_f4(get3(op.localAnchorB), op.wN);
				pts[2][p] = v3_f4(get3(op.angJA), op.lambdaN);
				pts[3][p] = v3_f4(get3(op.angJB), op.lambdaT);
				pts[4][p] = v3_f4(get3(op.rA), op.separation);
				pts[5][p] = v3_f4(get3(op.rB), u2f(op.id));
				pts[6][p] = v3_f4(get3(op.frictionAnchorA), 0.0f);
				pts[7][p] = v3_f4(get3(op.frictionAnchorB), 0.0f);
				pts[8][p] = v3_f4(get3(op.slipDir), 0.0f);
			}
		}
		DBuf<float4>* pcols[9] = { &w.cc.pA, &w.cc.pB, &w.cc.pJA, &w.cc.pJB, &w.cc.pRA, &w.cc.pRB, &w.cc.pFA,
								   &w.cc.pFB, &w.cc.pSlip };
		bool ok = hip_ok(hipMemcpy(w.cc.bodies.p, bodies.data(), C * sizeof(int2), hipMemcpyHostToDevice), "import") &&
				  hip_ok(hipMemcpy(w.cc.mat.p, mat.data(), C * sizeof(float4), hipMemcpyHostToDevice), "import") &&
				  hip_ok(hipMemcpy(w.cc.normal.p, nrm.data(), C * sizeof(float4), hipMemcpyHostToDevice), "import") &&
				  hip_ok(hipMemcpy(w.cc.cache.p, cch.data(), 4 * (size_t)C * sizeof(float4), hipMemcpyHostToDevice),
						 "import");
		for (int k = 0; k < 9 && ok; ++k)
			ok = hip_ok(hipMemcpy(pcols[k]->p, pts[k].data(), (size_t)4 * cap * sizeof(float4), hipMemcpyHostToDevice),
						"import");
		if (!ok)
		{
			m3d_world_destroy(world);
			return nullptr;
		}
	}
	w.contactCount = C;
	// stable tier
	w.stableCount = h.stableEntryCount;
	if (h.stableEntryCount > 0)
	{
		std::vector<uint64_t> k(h.stableEntryCount);
		std::vector<uint32_t> b(h.stableEntryCount);
		for (int32_t i = 0; i < h.stableEntryCount; ++i)
		{
			k[i] = s->stableEntries[i].key;
			b[i] = (uint32_t)s->stableEntries[i].body;
		}
		if (!w.stableKeys.ensure(k.size()) || !w.stableBodies.ensure(b.size()) ||
			!hip_ok(hipMemcpy(w.stableKeys.p, k.data(), k.size() * 8, hipMemcpyHostToDevice), "import") ||
			!hip_ok(hipMemcpy(w.stableBodies.p, b.data(), b.size() * 4, hipMemcpyHostToDevice), "import"))
		{
			m3d_world_destroy(world);
			return nullptr;
		}
	}
	w.stableOversized.assign(s->stableOversized, s->stableOversized + h.stableOversizedCount);
	if (!w.stableOversized.empty())
	{
		if (!w.dStableOversized.ensure(w.stableOversized.size()) ||
			!hip_ok(hipMemcpy(w.dStableOversized.p, w.stableOversized.data(), w.stableOversized.size() * 4,
							  hipMemcpyHostToDevice),
					"import"))
		{
			m3d_world_destroy(world);
			return nullptr;
		}
	}
	return world;
}

// ------------------------------------------------------------ views

int m3dx_debug_grid_entries(m3d_world* world, int tier, m3dx_grid_entry* out, int cap)
{
	World& w = world->w;
	int64_t n = tier == 0 ? w.activeCount : w.stableCount;
	if (out != nullptr && cap > 0 && n > 0)
	{
		int64_t m = std::min<int64_t>(n, cap);
		std::vector<uint64_t> k(m);
		std::vector<uint32_t> b(m);
		(void)hipMemcpy(k.data(), tier == 0 ? w.activeKeys.p : w.stableKeys.p, m * 8, hipMemcpyDeviceToHost);
		(void)hipMemcpy(b.data(), tier == 0 ? w.activeBodies.p : w.stableBodies.p, m * 4, hipMemcpyDeviceToHost);
		for (int64_t i = 0; i < m; ++i)
		{
			out[i].key = k[i];
			out[i].body = (int32_t)b[i];
			out[i].pad = 0;
		}
	}
	return (int)n;
}

int m3dx_debug_oversized(m3d_world* world, int tier, int32_t* out, int cap)
{
	World& w = world->w;
	const std::vector<int32_t>& v = tier == 0 ? w.activeOversized : w.stableOversized;
	for (int i = 0; i < (int)v.size() && i < cap && out != nullptr; ++i)
		out[i] = v[i];
	return (int)v.size();
}

int m3dx_debug_islands(m3d_world* world, int32_t* islandOfBody, int bodyCap, uint8_t* islandAwake, int islandCap)
{
	World& w = world->w;
	int32_t n = std::min<int32_t>(w.deviceBodyCount, bodyCap);
	if (!w.lastHadIslands || w.lastIslandCount == 0)
	{
		for (int i = 0; i < bodyCap && islandOfBody != nullptr; ++i)
			islandOfBody[i] = -1;
		return 0;
	}
	if (islandOfBody != nullptr)
	{
		for (int i = 0; i < bodyCap; ++i)
			islandOfBody[i] = -1;
		if (n > 0)
			(void)hipMemcpy(islandOfBody, w.islandOf.p, n * sizeof(int32_t), hipMemcpyDeviceToHost);
	}
	if (islandAwake != nullptr)
	{
		int32_t m = std::min<int32_t>(w.lastIslandCount, islandCap);
		std::vector<uint32_t> a(m);
		if (m > 0)
			(void)hipMemcpy(a.data(), w.islandAwakeD.p, m * sizeof(uint32_t), hipMemcpyDeviceToHost);
		for (int32_t i = 0; i < m; ++i)
			islandAwake[i] = a[i] ? 1 : 0;
	}
	return w.lastIslandCount;
}

int m3dx_debug_contact_colors(m3d_world* world, int32_t* colorOfContact, int cap)
{
	World& w = world->w;
	int32_t C = w.contactCount;
	int32_t m = std::min<int32_t>(C, cap);
	for (int i = 0; i < m; ++i)
		colorOfContact[i] = -1;
	if (w.lastHadIslands && m > 0)
		(void)hipMemcpy(colorOfContact, w.contactColor.p, m * sizeof(int32_t), hipMemcpyDeviceToHost);
	return C;
}

void m3dx_world_stats(m3d_world* world, m3dx_step_stats* out) { *out = world->w.stats; }

int m3dx_set_pass_blocks(m3d_world* world, int blocks)
{
	if (blocks < 0)
	{
		set_error("pass workgroups must be >= 0");
		return -1;
	}
	world->w.tuner.forcedBlocks = blocks;
	return 0;
}

int m3dx_pass_blocks(m3d_world* world) { return world->w.tuner.lastBlocks; }

int32_t m3dx_step_failures(m3d_world* world) { return world != nullptr ? world->w.stepFailures : -1; }

int m3dx_set_step_graphs(m3d_world* world, int enable)
{
	if (world == nullptr)
		return -1;
	world->w.stepGraphs = enable != 0;
	return 0;
}

int64_t m3dx_step_graph_launches(m3d_world* world) { return world != nullptr ? world->w.stepGraphLaunches : -1; }

int m3dx_set_xcd_barrier(m3d_world* world, int enable)
{
	if (world == nullptr)
		return -1;
	world->w.xcdBarrier = enable != 0;
	return 0;
}

int m3dx_set_grid_launch(m3d_world* world, int cooperative, uint32_t waitTicks)
{
	if (world == nullptr)
		return -1;
	world->w.coopLaunch = cooperative != 0;
	if (waitTicks > 0)
		world->w.barTimeout = waitTicks;
	return 0;
}

int m3dx_grid_launch_mode(m3d_world* world)
{
	if (world == nullptr)
		return -1;
	return world->w.coopLaunch ? 1 : 0;
}

int m3dx_set_kernel_timing(m3d_world* world, int enable)
{
	if (world == nullptr)
		return -1;
	World& w = world->w;
	w.timing = enable != 0;
	w.marks.clear();
	w.eventsUsed = 0;
	for (auto& t : w.ktime)
		t = m3dx_kernel_time{};
	return 0;
}

int m3dx_kernel_timing(m3d_world* world, m3dx_kernel_time* out)
{
	if (world == nullptr || out == nullptr)
		return -1;
	for (int k = 0; k < M3DX_K_COUNT; ++k)
		out[k] = world->w.ktime[k];
	return M3DX_K_COUNT;
}

} // extern "C"
