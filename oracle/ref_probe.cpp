// Test-harness shim compiled together with the reference engine's own
// sources (see oracle/Makefile). It exposes the reference world's internal
// state through the same m3dx_* snapshot/debug entry points that
// libmelon3d_amd.so exports, so the parity tests drive both engines through
// one binding. It is part of the oracle: only tests and the benchmark's CPU
// baseline load the library it is built into.
//
// Field meanings follow src/m3d_internal.h:77-229 (Body, Manifold, Contact),
// :313-349 (Joint), :372-497 (SpatialGrid) and :602-685 (m3d_world).

#define private public // reach SpatialGrid's tier storage for import/export
#include "m3d_internal.h"
#undef private

#include "melon3d_ext.h"

#include <string.h>

using namespace m3d;


namespace
{

void PutV3(float* d, m3d_vec3 v)
{
	d[0] = v.x;
	d[1] = v.y;
	d[2] = v.z;
}
void PutQ(float* d, m3d_quat q)
{
	d[0] = q.x;
	d[1] = q.y;
	d[2] = q.z;
	d[3] = q.w;
}
m3d_vec3 GetV3(const float* s) { return m3d_v3(s[0], s[1], s[2]); }
m3d_quat GetQ(const float* s)
{
	m3d_quat q = { s[0], s[1], s[2], s[3] };
	return q;
}

thread_local char g_err[256] = "";

} // namespace

extern "C"
{

const char* m3dx_last_error(void) { return g_err; }

int m3dx_set_device(int) { return 0; }

int m3dx_world_export(m3d_world* w, m3dx_snapshot* s)
{
	if (w == nullptr || s == nullptr)
	{
		return -1;
	}
	m3dx_world_header& h = s->header;
	memset(&h, 0, sizeof(h));
	const int32_t n = (int32_t)w->bodies.size();
	h.bodyCapacity = n;
	h.contactCount = (int32_t)w->contacts.size();
	h.jointCapacity = (int32_t)w->joints.size();
	h.freeBodyCount = (int32_t)w->freeBodies.size();
	h.freeJointCount = (int32_t)w->freeJoints.size();
	h.stableEntryCount = (int32_t)w->grid.m_stable.size();
	h.stableOversizedCount = (int32_t)w->grid.m_stableOversized.size();
	h.gridDirty = w->gridDirty;
	h.stableDirty = w->stableDirty;
	h.stableForce = w->stableForce;
	h.stepsSinceStableBuild = w->stepsSinceStableBuild;
	h.anyForcesApplied = w->anyForcesApplied;
	h.jointedPairsDirty = w->jointedPairsDirty;
	PutV3(h.gravity, w->def.gravity);
	h.workerCount = w->def.workerCount;
	h.positionIterations = w->def.positionIterations;
	h.gridCellSize = w->def.gridCellSize;
	h.enableSleep = w->def.enableSleep;

	if (s->bodies != nullptr)
	{
		for (int32_t i = 0; i < n; ++i)
		{
			const Body& b = w->bodies[i];
			m3dx_body_state& o = s->bodies[i];
			memset(&o, 0, sizeof(o));
			PutV3(o.position, b.position);
			PutQ(o.rotation, b.rotation);
			PutV3(o.linearVelocity, b.linearVelocity);
			PutV3(o.angularVelocity, b.angularVelocity);
			PutV3(o.prevPosition, b.prevPosition);
			PutQ(o.prevRotation, b.prevRotation);
			PutV3(o.prevLinearVelocity, b.prevLinearVelocity);
			PutV3(o.prevAngularVelocity, b.prevAngularVelocity);
			o.invMass = b.invMass;
			PutV3(o.invInertiaLocal, b.invInertiaLocal);
			PutV3(o.invInertiaWorld + 0, b.invInertiaWorld.cx);
			PutV3(o.invInertiaWorld + 3, b.invInertiaWorld.cy);
			PutV3(o.invInertiaWorld + 6, b.invInertiaWorld.cz);
			o.shapeType = (int32_t)b.shape.type;
			o.radius = b.shape.radius;
			o.halfHeight = b.shape.halfHeight;
			PutV3(o.halfExtents, b.shape.halfExtents);
			o.friction = b.friction;
			o.restitution = b.restitution;
			o.rollingResistance = b.rollingResistance;
			o.mass = b.mass;
			PutV3(o.force, b.force);
			PutV3(o.torque, b.torque);
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
			if (i < (int32_t)w->expandedAABBs.size())
			{
				PutV3(o.fatAABB, w->expandedAABBs[i].lowerBound);
				PutV3(o.fatAABB + 3, w->expandedAABBs[i].upperBound);
			}
			o.hasFatAABB = i < (int32_t)w->hasFatAABB.size() ? w->hasFatAABB[i] : 0;
			o.inStableTier = i < (int32_t)w->inStableTier.size() ? w->inStableTier[i] : 0;
			o.inStableCells = w->grid.InStableCells(i) ? 1 : 0;
		}
	}
	if (s->contacts != nullptr)
	{
		for (int32_t i = 0; i < h.contactCount; ++i)
		{
			const Contact& c = w->contacts[i];
			m3dx_contact_state& o = s->contacts[i];
			memset(&o, 0, sizeof(o));
			o.bodyA = c.bodyA;
			o.bodyB = c.bodyB;
			o.friction = c.friction;
			o.restitution = c.restitution;
			o.rollingResistance = c.rollingResistance;
			o.touching = c.touching;
			o.isSensor = c.isSensor;
			o.tunnelGuard = c.tunnelGuard;
			o.curved = c.curved;
			o.hasCache = c.hasCache;
			PutV3(o.cachePosA, c.cachePosA);
			PutV3(o.cachePosB, c.cachePosB);
			PutQ(o.cacheRotA, c.cacheRotA);
			PutQ(o.cacheRotB, c.cacheRotB);
			PutV3(o.normal, c.manifold.normal);
			o.pointCount = c.manifold.pointCount;
			for (int p = 0; p < 4; ++p)
			{
				const ManifoldPoint& mp = c.manifold.points[p];
				m3dx_point_state& op = o.points[p];
				PutV3(op.localAnchorA, mp.localAnchorA);
				PutV3(op.localAnchorB, mp.localAnchorB);
				PutV3(op.frictionAnchorA, mp.frictionAnchorA);
				PutV3(op.frictionAnchorB, mp.frictionAnchorB);
				op.separation = mp.separation;
				op.id = mp.id;
				op.lambdaN = mp.lambdaN;
				op.lambdaT = mp.lambdaT;
				op.separationOffset = mp.separationOffset;
				PutV3(op.slipDir, mp.slipDir);
				PutV3(op.rA, mp.rA);
				PutV3(op.rB, mp.rB);
				op.wN = mp.wN;
				PutV3(op.angJA, mp.angJA);
				PutV3(op.angJB, mp.angJB);
			}
		}
	}
	if (s->joints != nullptr)
	{
		for (int32_t i = 0; i < h.jointCapacity; ++i)
		{
			const Joint& j = w->joints[i];
			m3dx_joint_state& o = s->joints[i];
			memset(&o, 0, sizeof(o));
			o.type = (int32_t)j.type;
			o.bodyA = j.bodyA;
			o.bodyB = j.bodyB;
			PutV3(o.localAnchorA, j.localAnchorA);
			PutV3(o.localAnchorB, j.localAnchorB);
			o.length = j.length;
			o.hertz = j.hertz;
			o.dampingRatio = j.dampingRatio;
			PutV3(o.localAxisA, j.localAxisA);
			PutV3(o.localAxisB, j.localAxisB);
			PutV3(o.localRefA, j.localRefA);
			PutV3(o.localRefB, j.localRefB);
			o.enableLimit = j.enableLimit;
			o.lowerAngle = j.lowerAngle;
			o.upperAngle = j.upperAngle;
			o.enableMotor = j.enableMotor;
			o.motorSpeed = j.motorSpeed;
			o.maxMotorTorque = j.maxMotorTorque;
			PutQ(o.relRot0, j.relRot0);
			o.compliance = j.compliance;
			o.lambda = j.lambda;
			o.motorLambda = j.motorLambda;
			o.inUse = j.inUse;
			o.generation = j.generation;
		}
	}
	if (s->freeBodies != nullptr)
	{
		for (int32_t i = 0; i < h.freeBodyCount; ++i)
		{
			s->freeBodies[i] = w->freeBodies[i];
		}
	}
	if (s->freeJoints != nullptr)
	{
		for (int32_t i = 0; i < h.freeJointCount; ++i)
		{
			s->freeJoints[i] = w->freeJoints[i];
		}
	}
	if (s->stableEntries != nullptr)
	{
		for (int32_t i = 0; i < h.stableEntryCount; ++i)
		{
			s->stableEntries[i].key = w->grid.m_stable[i].key;
			s->stableEntries[i].body = w->grid.m_stable[i].body;
			s->stableEntries[i].pad = 0;
		}
	}
	if (s->stableOversized != nullptr)
	{
		for (int32_t i = 0; i < h.stableOversizedCount; ++i)
		{
			s->stableOversized[i] = w->grid.m_stableOversized[i];
		}
	}
	return 0;
}

m3d_world* m3dx_world_import(const m3dx_snapshot* s)
{
	if (s == nullptr)
	{
		return nullptr;
	}
	const m3dx_world_header& h = s->header;
	m3d_world_def def = m3d_world_def_default();
	def.gravity = GetV3(h.gravity);
	def.workerCount = h.workerCount;
	def.positionIterations = h.positionIterations;
	def.gridCellSize = h.gridCellSize;
	def.enableSleep = h.enableSleep != 0;
	m3d_world* w = m3d_world_create(&def);
	if (w == nullptr)
	{
		return nullptr;
	}
	const int32_t n = h.bodyCapacity;
	w->bodies.resize(n);
	w->expandedAABBs.resize(n);
	w->hasFatAABB.assign(n, 0);
	w->inStableTier.assign(n, 0);
	w->grid.m_inStable.assign(n, false);
	for (int32_t i = 0; i < n; ++i)
	{
		const m3dx_body_state& o = s->bodies[i];
		Body& b = w->bodies[i];
		b = Body();
		b.position = GetV3(o.position);
		b.rotation = GetQ(o.rotation);
		b.linearVelocity = GetV3(o.linearVelocity);
		b.angularVelocity = GetV3(o.angularVelocity);
		b.prevPosition = GetV3(o.prevPosition);
		b.prevRotation = GetQ(o.prevRotation);
		b.prevLinearVelocity = GetV3(o.prevLinearVelocity);
		b.prevAngularVelocity = GetV3(o.prevAngularVelocity);
		b.invMass = o.invMass;
		b.invInertiaLocal = GetV3(o.invInertiaLocal);
		b.invInertiaWorld.cx = GetV3(o.invInertiaWorld + 0);
		b.invInertiaWorld.cy = GetV3(o.invInertiaWorld + 3);
		b.invInertiaWorld.cz = GetV3(o.invInertiaWorld + 6);
		b.shape.type = (m3d_shape_type)o.shapeType;
		b.shape.radius = o.radius;
		b.shape.halfHeight = o.halfHeight;
		b.shape.halfExtents = GetV3(o.halfExtents);
		b.friction = o.friction;
		b.restitution = o.restitution;
		b.rollingResistance = o.rollingResistance;
		b.mass = o.mass;
		b.force = GetV3(o.force);
		b.torque = GetV3(o.torque);
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
		w->expandedAABBs[i].lowerBound = GetV3(o.fatAABB);
		w->expandedAABBs[i].upperBound = GetV3(o.fatAABB + 3);
		w->hasFatAABB[i] = (uint8_t)(o.hasFatAABB != 0);
		w->inStableTier[i] = (uint8_t)(o.inStableTier != 0);
		w->grid.m_inStable[i] = o.inStableCells != 0;
	}
	w->freeBodies.assign(s->freeBodies, s->freeBodies + h.freeBodyCount);
	w->joints.resize(h.jointCapacity);
	for (int32_t i = 0; i < h.jointCapacity; ++i)
	{
		const m3dx_joint_state& o = s->joints[i];
		Joint& j = w->joints[i];
		j = Joint();
		j.type = (JointType)o.type;
		j.bodyA = o.bodyA;
		j.bodyB = o.bodyB;
		j.localAnchorA = GetV3(o.localAnchorA);
		j.localAnchorB = GetV3(o.localAnchorB);
		j.length = o.length;
		j.hertz = o.hertz;
		j.dampingRatio = o.dampingRatio;
		j.localAxisA = GetV3(o.localAxisA);
		j.localAxisB = GetV3(o.localAxisB);
		j.localRefA = GetV3(o.localRefA);
		j.localRefB = GetV3(o.localRefB);
		j.enableLimit = o.enableLimit != 0;
		j.lowerAngle = o.lowerAngle;
		j.upperAngle = o.upperAngle;
		j.enableMotor = o.enableMotor != 0;
		j.motorSpeed = o.motorSpeed;
		j.maxMotorTorque = o.maxMotorTorque;
		j.relRot0 = GetQ(o.relRot0);
		j.compliance = o.compliance;
		j.lambda = o.lambda;
		j.motorLambda = o.motorLambda;
		j.inUse = o.inUse != 0;
		j.generation = o.generation;
	}
	w->freeJoints.assign(s->freeJoints, s->freeJoints + h.freeJointCount);
	for (int32_t i = 0; i < h.contactCount; ++i)
	{
		const m3dx_contact_state& o = s->contacts[i];
		Contact c{};
		c.bodyA = o.bodyA;
		c.bodyB = o.bodyB;
		c.key = ContactKey(o.bodyA, o.bodyB);
		c.friction = o.friction;
		c.restitution = o.restitution;
		c.rollingResistance = o.rollingResistance;
		c.touching = o.touching != 0;
		c.isSensor = o.isSensor != 0;
		c.tunnelGuard = o.tunnelGuard != 0;
		c.curved = o.curved != 0;
		c.hasCache = o.hasCache != 0;
		c.cachePosA = GetV3(o.cachePosA);
		c.cachePosB = GetV3(o.cachePosB);
		c.cacheRotA = GetQ(o.cacheRotA);
		c.cacheRotB = GetQ(o.cacheRotB);
		c.manifold.normal = GetV3(o.normal);
		c.manifold.pointCount = o.pointCount;
		for (int p = 0; p < 4; ++p)
		{
			const m3dx_point_state& op = o.points[p];
			ManifoldPoint& mp = c.manifold.points[p];
			mp.localAnchorA = GetV3(op.localAnchorA);
			mp.localAnchorB = GetV3(op.localAnchorB);
			mp.frictionAnchorA = GetV3(op.frictionAnchorA);
			mp.frictionAnchorB = GetV3(op.frictionAnchorB);
			mp.separation = op.separation;
			mp.id = op.id;
			mp.lambdaN = op.lambdaN;
			mp.lambdaT = op.lambdaT;
			mp.separationOffset = op.separationOffset;
			mp.slipDir = GetV3(op.slipDir);
			mp.rA = GetV3(op.rA);
			mp.rB = GetV3(op.rB);
			mp.wN = op.wN;
			mp.angJA = GetV3(op.angJA);
			mp.angJB = GetV3(op.angJB);
		}
		w->contactMap[c.key] = i;
		w->contacts.push_back(c);
		w->cBodyA.push_back(c.bodyA);
		w->cBodyB.push_back(c.bodyB);
		w->cTouching.push_back(c.touching ? 1 : 0);
		w->cSensor.push_back(c.isSensor ? 1 : 0);
	}
	w->grid.m_cellSize = h.gridCellSize;
	w->grid.m_stable.resize(h.stableEntryCount);
	for (int32_t i = 0; i < h.stableEntryCount; ++i)
	{
		w->grid.m_stable[i].key = s->stableEntries[i].key;
		w->grid.m_stable[i].body = s->stableEntries[i].body;
	}
	w->grid.m_stableOversized.assign(s->stableOversized, s->stableOversized + h.stableOversizedCount);
	w->gridDirty = h.gridDirty != 0;
	w->stableDirty = h.stableDirty != 0;
	w->stableForce = h.stableForce != 0;
	w->stepsSinceStableBuild = h.stepsSinceStableBuild;
	w->anyForcesApplied = h.anyForcesApplied != 0;
	w->jointedPairsDirty = true;
	return w;
}

int m3dx_debug_grid_entries(m3d_world* w, int tier, m3dx_grid_entry* out, int cap)
{
	const auto& e = tier == 0 ? w->grid.m_active : w->grid.m_stable;
	int n = (int)e.size();
	for (int i = 0; i < n && i < cap && out != nullptr; ++i)
	{
		out[i].key = e[i].key;
		out[i].body = e[i].body;
		out[i].pad = 0;
	}
	return n;
}

int m3dx_debug_oversized(m3d_world* w, int tier, int32_t* out, int cap)
{
	const auto& e = tier == 0 ? w->grid.m_activeOversized : w->grid.m_stableOversized;
	int n = (int)e.size();
	for (int i = 0; i < n && i < cap && out != nullptr; ++i)
	{
		out[i] = e[i];
	}
	return n;
}

int m3dx_debug_islands(m3d_world* w, int32_t* islandOfBody, int bodyCap, uint8_t* islandAwake, int islandCap)
{
	if (w->profile.islandCount == 0)
	{
		for (int i = 0; i < bodyCap && islandOfBody != nullptr; ++i)
		{
			islandOfBody[i] = -1;
		}
		return 0;
	}
	int count = w->islandUsed;
	for (int i = 0; i < bodyCap && islandOfBody != nullptr; ++i)
	{
		islandOfBody[i] = -1;
	}
	for (int idx = 0; idx < count; ++idx)
	{
		for (int32_t bi : w->islandBodies[idx])
		{
			if (bi < bodyCap && islandOfBody != nullptr)
			{
				islandOfBody[bi] = idx;
			}
		}
		if (islandAwake != nullptr && idx < islandCap)
		{
			islandAwake[idx] = w->islandAwake[idx];
		}
	}
	return count;
}

// Only the most recently colored island's buckets survive a reference step
// (the scratch is reused per colored island), so this view is meaningful for
// scenes with at most one colored island.
int m3dx_debug_contact_colors(m3d_world* w, int32_t* colorOfContact, int cap)
{
	int n = (int)w->contacts.size();
	for (int i = 0; i < cap && i < n; ++i)
	{
		colorOfContact[i] = -1;
	}
	for (int c = 0; c < (int)w->colorBuckets.size() && c < kMaxGraphColors; ++c)
	{
		for (int32_t ci : w->colorBuckets[c])
		{
			if (ci < cap)
			{
				colorOfContact[ci] = c;
			}
		}
	}
	for (int32_t ci : w->overflowContacts)
	{
		if (ci < cap)
		{
			colorOfContact[ci] = kMaxGraphColors;
		}
	}
	return n;
}

void m3dx_world_stats(m3d_world* w, m3dx_step_stats* out)
{
	memset(out, 0, sizeof(*out));
	out->bodyCapacity = (int32_t)w->bodies.size();
	out->contacts = (int32_t)w->contacts.size();
	out->islands = w->profile.islandCount;
	out->activeEntries = (int32_t)w->grid.m_active.size();
	out->stableEntries = (int32_t)w->grid.m_stable.size();
}

/* batched worlds are a device-backend extension: the CPU engine steps one
   world per call, so a batch is a loop over plain worlds there */
m3d_world* m3dx_batch_create(const m3d_world_def*, int32_t) { return nullptr; }
int m3dx_batch_select(m3d_world*, int32_t) { return -1; }
int32_t m3dx_batch_size(m3d_world*) { return 0; }
int m3dx_batch_body_range(m3d_world*, int32_t, int32_t*, int32_t*) { return -1; }
int m3dx_batch_contact_counts(m3d_world*, int32_t*, int32_t) { return 0; }

/* the adjoint replays recorded device steps: not available on the CPU engine */
int m3dx_diff_record(m3d_world*, int) { return -1; }
int32_t m3dx_diff_steps(m3d_world*) { return 0; }
int m3dx_diff_state(m3d_world*, int, float*) { return -1; }
int m3dx_diff_replay(m3d_world*, int, int, int, const void*, const void*, const void*, void*) { return -1; }
int m3dx_diff_backward(m3d_world*, const float*, float*, float*, float*) { return -1; }
int m3dx_diff_backward_tape(m3d_world*, const float*, float*, float*, float*) { return -1; }
int m3dx_diff_backward_f64(m3d_world*, const double*, double*, double*, double*) { return -1; }
int m3dx_diff_set_checkpoint(m3d_world*, int) { return -1; }
int m3dx_diff_forward(m3d_world*, int, int, float*) { return -1; }
int m3dx_diff_jvp_cpu(m3d_world*, int, int, int, const double*, const double*, const double*, const double*, const double*,
                      const double*, double*, double*)
{
    return -1;
}

/* no device kernels on the CPU engine: timing is unavailable */
int m3dx_set_kernel_timing(m3d_world*, int) { return -1; }
int m3dx_set_pass_blocks(m3d_world*, int) { return -1; }
int m3dx_world_ray_cast_batch(m3d_world*, const m3d_vec3*, const m3d_vec3*, const uint32_t*, int32_t, m3d_ray_result*)
{
    return -1;
}
/* the CPU engine's overlap queries are its callback query, one per box */
int m3dx_world_overlap_aabb_batch(m3d_world* w, const m3d_aabb* boxes, int32_t count, int32_t* offsets,
                                  m3d_body_id* out, int32_t cap)
{
    if (w == nullptr || count < 0 || (count > 0 && (boxes == nullptr || offsets == nullptr)) || cap < 0 ||
        (cap > 0 && out == nullptr))
        return -1;
    struct Sink
    {
        m3d_body_id* out;
        int32_t cap, n;
    } sink = { out, cap, 0 };
    for (int32_t k = 0; k < count; ++k)
    {
        offsets[k] = sink.n;
        m3d_world_overlap_aabb(
            w, boxes[k],
            [](m3d_body_id id, void* ctx) {
                Sink* s = (Sink*)ctx;
                if (s->n < s->cap)
                    s->out[s->n] = id;
                ++s->n;
                return true;
            },
            &sink);
    }
    offsets[count] = sink.n;
    return sink.n;
}
/* the device sort hook has no CPU engine counterpart */
int m3dx_sort_pairs(const uint64_t*, const uint32_t*, int64_t, int, uint64_t*, uint32_t*)
{
    return -1;
}
/* the CPU engine's transforms are its getter, one body at a time */
int m3dx_world_transforms(m3d_world* w, const m3d_body_id* ids, int32_t count, m3d_transform* out)
{
    if (w == nullptr || count < 0 || (count > 0 && (ids == nullptr || out == nullptr)))
        return -1;
    for (int32_t k = 0; k < count; ++k)
        out[k] = m3d_body_transform(w, ids[k]);
    return 0;
}
/* the bulk spawn/despawn is defined as the per-call sequence: on the CPU
   engine it is that sequence */
int m3dx_body_respawn(m3d_world* w, const m3d_body_id* destroy, const m3d_body_def* bodyDefs,
                      const m3d_shape_def* shapeDefs, int32_t count, m3d_body_id* ids)
{
    if (w == nullptr || count < 0 || (bodyDefs == nullptr) != (shapeDefs == nullptr) ||
        (bodyDefs != nullptr && ids == nullptr))
        return -1;
    for (int32_t k = 0; k < count; ++k)
    {
        if (destroy != nullptr)
            m3d_body_destroy(w, destroy[k]);
        if (bodyDefs != nullptr)
            ids[k] = m3d_body_create(w, &bodyDefs[k], &shapeDefs[k]);
    }
    return 0;
}
int m3dx_world_ray_cast_batch_noalloc(m3d_world*, const m3d_vec3*, const m3d_vec3*, const uint32_t*, int32_t,
                                      m3d_ray_result*)
{
    return -1;
}
int m3dx_world_sphere_cast_batch(m3d_world*, const m3d_vec3*, const float*, const m3d_vec3*, const uint32_t*, int32_t,
                                 m3d_ray_result*)
{
    return -1;
}
int m3dx_pass_blocks(m3d_world*) { return 0; }
int32_t m3dx_step_failures(m3d_world*) { return 0; }
int m3dx_set_step_graphs(m3d_world*, int) { return -1; }
int64_t m3dx_step_graph_launches(m3d_world*) { return 0; }
int m3dx_set_xcd_barrier(m3d_world*, int) { return -1; }
int m3dx_set_grid_launch(m3d_world*, int, uint32_t) { return -1; }
int m3dx_grid_launch_mode(m3d_world*) { return -1; }

int m3dx_kernel_timing(m3d_world*, m3dx_kernel_time* out)
{
	if (out != nullptr)
		memset(out, 0, sizeof(m3dx_kernel_time) * M3DX_K_COUNT);
	return M3DX_K_COUNT;
}

} // extern "C"
