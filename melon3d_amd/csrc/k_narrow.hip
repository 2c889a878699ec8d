// Step narrowphase: one thread per contact (src/world.cpp:1738-1815), then
// touch transitions in contact order: begin/end events and the
// wake-on-touch pass (src/world.cpp:1817-1845).
//
// The reference walks transitions serially and wakes a sleeping partner of
// an awake body, so a body woken by an earlier transition can wake another
// at a later one. The device computes the same result as the least fixpoint
// of wakeAt[x] = min(initial, t : transition t touches x and its partner was
// awake before t), iterated until no value changes.
#include "mx_steps.h"
#include "mx_wave.h"

namespace mx
{

DBodies make_body_view(World& w);
DContacts make_contact_view(World& w);

constexpr int kBlockN = 256;
static inline unsigned gridn(int64_t n) { return (unsigned)((n + kBlockN - 1) / kBlockN); }

// One contact's narrowphase; true when its manifold was regenerated.
__device__ __forceinline__ bool narrow_contact(const DBodies& B, const DContacts& C, int32_t c, float dt, int8_t* trans)
{
	int2 ab = C.bodies[c];
	int32_t ia = ab.x, ib = ab.y;
	if (!(B.sb[ia] & SB_AWAKE) && !(B.sb[ib] & SB_AWAKE))
	{
		trans[c] = 0;
		return false;
	}
	float4 mat = C.mat[c];
	uint32_t cf = f2u(mat.w);
	m3d_vec3 pa = f4_v3(B.pos[ia]), pb = f4_v3(B.pos[ib]);
	m3d_quat qa = f4_q(B.rot[ia]), qb = f4_q(B.rot[ib]);
	int32_t pointCount = f2i(C.normal[c].w);

	if (cf & CF_TUNNEL_GUARD)
	{
		if (tracked_min_sep(pa, qa, pb, qb, StoredManifold(C, c)) < -kLinearSlop)
		{
			trans[c] = 0;
			return false;
		}
		cf &= ~CF_TUNNEL_GUARD;
	}

	bool regen = !(cf & CF_HAS_CACHE);
	if (!regen)
	{
		const float4* k = cache_of(C, c);
		regen = pose_changed(pa, qa, f4_v3(k[0]), f4_q(k[1])) || pose_changed(pb, qb, f4_v3(k[2]), f4_q(k[3]));
	}
	if (regen)
	{
		// the old manifold is read where it lies; the new one is built here
		// (slots beyond its count are never stored and keep their contents)
		const StoredManifold old(C, c);
		Manifold m;
		Shape sA = load_shape(B, ia), sB = load_shape(B, ib);
		float relSpeed = m3d_length(m3d_sub(f4_v3(B.vel[ia]), f4_v3(B.vel[ib])));
		float margin = kSpeculativeBase + dt * relSpeed;
		m3d_transform xfA = { pa, qa }, xfB = { pb, qb };
		collide_shapes(sA, xfA, sB, xfB, margin, &m);
		inherit_friction_anchors(old, m, pa, qa, pb, qb, sA.type == M3D_SHAPE_BOX && sB.type == M3D_SHAPE_BOX);
		cf |= CF_HAS_CACHE;
		float4* k = cache_of(C, c);
		k[0] = v3_f4(pa, 0.0f);
		k[1] = q_f4(qa);
		k[2] = v3_f4(pb, 0.0f);
		k[3] = q_f4(qb);
		set_separation_offsets(m, pa, qa, pb, qb);
		store_manifold(C, c, m);
		pointCount = m.pointCount;
	}
	bool was = (cf & CF_TOUCHING) != 0;
	bool now = pointCount > 0;
	cf = now ? (cf | CF_TOUCHING) : (cf & ~CF_TOUCHING);
	trans[c] = (int8_t)(now && !was ? 1 : (!now && was ? -1 : 0));
	C.mat[c] = make_float4(mat.x, mat.y, mat.z, u2f(cf));
	return regen;
}

// The regeneration count is summed per workgroup (one atomic each): an
// atomic from every wave on one counter serializes.
__global__ void __launch_bounds__(256) k_narrow(DBodies B, DContacts C, int32_t count, float dt, int8_t* trans,
												uint32_t* counters)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	bool regen = c < count && narrow_contact(B, C, c, dt, trans);
	block_add<1>(counters, { CT_REGEN }, { regen ? 1u : 0u });
}

__global__ void k_trans_flags(const int8_t* trans, int32_t count, uint32_t* flags)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c < count)
		flags[c] = trans[c] != 0 ? 1u : 0u;
}

// The transitions in contact order: list[t] = contact, out[t] = (body a,
// body b, kind, contact) with kind +-1 (begin / end), +-2 for sensors.
__global__ void k_trans_compact(const int2* bodies, const float4* mat, const int8_t* trans, const uint32_t* offs,
								int32_t count, int32_t* list, int4* out)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= count || trans[c] == 0)
		return;
	const uint32_t t = offs[c];
	list[t] = c;
	int2 b = bodies[c];
	uint32_t cf = f2u(mat[c].w);
	out[t] = make_int4(b.x, b.y, (int)trans[c] * (((cf & CF_SENSOR) != 0) ? 2 : 1), c);
}

__global__ void k_wake_init(const uint32_t* flags, int32_t n, int32_t* wakeAt)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n)
		wakeAt[i] = (flags[i] & BF_AWAKE) ? -1 : 0x7FFFFFFF;
}

// info.z: +1 begin, +2 begin (sensor), -1/-2 end
// The wake-on-touch fixpoint in one workgroup: a touching begin t (info.z =
// 1) with one body awake before t (wake time below t) wakes the other body
// at t (its wake time lowered to t; static bodies stay). Rounds over every
// transition until one changes nothing: a wake time only decreases, so the
// rounds are bounded by the longest wake chain (T + 2 caps them). Wake times
// are read and lowered at agent scope.
__global__ void __launch_bounds__(1024) k_wake_fixpoint(const int4* info, const uint32_t* count, const uint32_t* flags,
														int32_t* wakeAt)
{
	__shared__ int changed;
	const int32_t T = (int32_t)*count;
	for (int32_t round = 0; round < T + 2; ++round)
	{
		if (threadIdx.x == 0)
			changed = 0;
		__syncthreads();
		bool mine = false;
		for (int32_t t = (int32_t)threadIdx.x; t < T; t += (int32_t)blockDim.x)
		{
			int4 e = info[t];
			if (e.z != 1)
				continue;
			int32_t a = e.x, b = e.y;
			int32_t wa = __hip_atomic_load(&wakeAt[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			int32_t wb = __hip_atomic_load(&wakeAt[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			bool awA = wa < t, awB = wb < t;
			int32_t other = awA && !awB ? b : (awB && !awA ? a : -1);
			if (other >= 0 && bf_type(flags[other]) != M3D_BODY_STATIC &&
				__hip_atomic_fetch_min(&wakeAt[other], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > t)
				mine = true;
		}
		if (mine)
			changed = 1;
		__syncthreads();
		bool done = changed == 0;
		__syncthreads();
		if (done)
			break;
	}
}

// WakeBodyInternal for bodies woken by a transition; src/world.cpp:121-140
__global__ void k_wake_apply(DBodies B, int32_t n, const int32_t* wakeAt, uint32_t* counters)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i >= n)
		return;
	int32_t wv = wakeAt[i];
	if (wv < 0 || wv == 0x7FFFFFFF)
		return;
	B.flags[i] |= BF_AWAKE;
	B.sb[i] |= SB_AWAKE;
	float4 a = B.iw1[i];
	a.w = 0.0f; // sleepTime
	B.iw1[i] = a;
	float4 b = B.iw2[i];
	b.w = 4.0f * kSleepEnergyThresh; // smoothedEnergy
	B.iw2[i] = b;
	counters[CT_STABLE_DIRTY] = 1u;
	mark_sub_stable_dirty(B, i);
}

bool narrowphase(World& w, float dt)
{
	hipStream_t s = w.stream;
	int32_t C = w.contactCount;
	int32_t n = w.deviceBodyCount;
	if (C == 0)
		return true;
	if (!w.u8a.ensure(C) || !w.u32a.ensure(C) || !w.u32b.ensure(C) || !w.i32a.ensure(C) ||
		!ensure_prim_scratch(w, C) || !w.transList.ensure(C) || !w.transInfo.ensure(C) || !w.i32c.ensure(n))
		return false;
	int8_t* trans = (int8_t*)w.u8a.p;
	w.mark_begin(M3DX_K_NARROW);
	k_narrow<<<gridn(C), kBlockN, 0, s>>>(make_body_view(w), make_contact_view(w), C, dt, trans, w.counters.p);
	w.mark_end();
	k_trans_flags<<<gridn(C), kBlockN, 0, s>>>(trans, C, w.u32a.p);
	scan_exclusive(w.u32a.p, w.u32b.p, C, w.counters.p + CT_TRANSITIONS, w.scanStatus.p, s);
	// The transitions are listed and described on the device (their count
	// read there); the count comes back with the island stage's first
	// readback, which issues their download (queue_transitions).
	k_trans_compact<<<gridn(C), kBlockN, 0, s>>>(w.cc.bodies.p, w.cc.mat.p, trans, w.u32b.p, C, w.transList.p,
												 w.transInfo.p);
	// wake-on-touch fixpoint
	int32_t* wakeAt = w.i32c.p;
	k_wake_init<<<gridn(n), kBlockN, 0, s>>>(w.bc.flags.p, n, wakeAt);
	k_wake_fixpoint<<<1, 1024, 0, s>>>(w.transInfo.p, w.counters.p + CT_TRANSITIONS, w.bc.flags.p, wakeAt);
	k_wake_apply<<<gridn(n), kBlockN, 0, s>>>(make_body_view(w), n, wakeAt, w.counters.p);
	w.transInFlight = true;
	return true;
}

// Download of the step's transitions once their count is in the host
// counters: it runs behind the launches queued so far, and the step turns
// them into events after the solver's launches are queued
// (flush_transitions).
bool queue_transitions(World& w)
{
	if (!w.transInFlight)
		return true;
	w.transInFlight = false;
	int32_t T = (int32_t)w.hostCounters[CT_TRANSITIONS];
	if (T == 0)
		return true;
	hipStream_t s = w.stream;
	if (w.pinnedTransCap < (size_t)T)
	{
		if (w.pinnedTrans != nullptr)
			(void)hipHostFree(w.pinnedTrans);
		w.pinnedTrans = nullptr;
		w.pinnedTransCap = (size_t)T * 3 / 2 + 256;
		if (!hip_ok(hipHostMalloc((void**)&w.pinnedTrans, w.pinnedTransCap * sizeof(int4), hipHostMallocDefault),
					"transition staging"))
		{
			w.pinnedTransCap = 0;
			return false;
		}
	}
	if (w.transReady == nullptr && !hip_ok(hipEventCreateWithFlags(&w.transReady, hipEventDisableTiming), "transitions"))
		return false;
	if (!hip_ok(hipMemcpyAsync(w.pinnedTrans, w.transInfo.p, T * sizeof(int4), hipMemcpyDeviceToHost, s),
				"transitions") ||
		!hip_ok(hipEventRecord(w.transReady, s), "transitions"))
		return false;
	w.pendingTrans = T;
	return true;
}

// Begin and end events of the step's touch transitions, in contact order,
// after the removal's end events (src/world.cpp:1773-1801).
bool flush_transitions(World& w)
{
	// a step that read no counters back after the narrowphase
	if (w.transInFlight && !(sync_counters(w) && queue_transitions(w)))
		return false;
	int32_t T = w.pendingTrans;
	if (T == 0)
		return true;
	w.pendingTrans = 0;
	if (!hip_ok(hipEventSynchronize(w.transReady), "transitions"))
		return false;
	for (int32_t t = 0; t < T; ++t)
	{
		const int4& e = w.pinnedTrans[t];
		m3d_contact_event ce;
		ce.bodyA = { e.x, w.bodies[e.x].generation };
		ce.bodyB = { e.y, w.bodies[e.y].generation };
		ce.isSensor = e.z == 2 || e.z == -2;
		if (e.z > 0)
			w.beginEvents.push_back(ce);
		else
			w.endEvents.push_back(ce);
	}
	return true;
}

} // namespace mx
