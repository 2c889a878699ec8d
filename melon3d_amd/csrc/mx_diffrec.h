// Recorded steps of a differentiable world: the records the engine writes
// while recording (k_diff.hip) and the views the replays and adjoints read
// (k_diff.hip, k_adjoint.hip).
#pragma once

#include "mx_diff.h"
#include "mx_steps.h"

#include <string>
#include <vector>

namespace mx
{

using dd::Ad;
using dd::StepIn;
using dd::StepState;
using dd::Tape;

constexpr int kLeaves = 15; // per body: x(3) q(4) v(3) w(3) mass friction
constexpr int kOuts = 13;   // per body: x(3) q(4) v(3) w(3)

struct DiffStepRec
{
	int32_t nb = 0, nc = 0;
	int32_t maxWorldBodies = 0, maxWorldPoints = 0;
	float dt = 0.0f;
	int substeps = 1;
	m3d_vec3 gravity{};
	DBuf<float4> pos, rot, vel, angVel, invIL, shape, bmat;
	DBuf<float> angDamp;
	DBuf<uint32_t> flags;
	DBuf<int32_t> contactStart; // worlds + 1
	DBuf<int2> bodies;          // world-local
	DBuf<int32_t> idx;          // engine contact slot of each recorded contact
	// one block per substep: the manifold the substep's prepare left
	DBuf<float4> normal, cmat;            // [s * nc + c]
	DBuf<float4> pA, pB, pFA, pFB, pSlip; // [(s * nc + c) * 4 + j]
	DBuf<int32_t> sched;                  // [s * nc + c]: schedule word (mx_diff.h)
	void release()
	{
		DBuf<float4>* f[] = { &pos, &rot, &vel, &angVel, &invIL, &shape, &bmat, &normal, &cmat, &pA, &pB, &pFA, &pFB, &pSlip };
		for (auto* p : f)
			p->release();
		angDamp.release();
		flags.release();
		contactStart.release();
		bodies.release();
		idx.release();
		sched.release();
	}
};

struct DiffState
{
	bool recording = false;
	// step at which recording was refused (-1: none): a step whose solve the
	// adjoint does not cover (joints) stops the
	// recording, the step itself runs as usual, and every later use of the
	// records fails with `refusal`
	int32_t refusedStep = -1;
	std::string refusal;
	int32_t worlds = 1;
	std::vector<int32_t> bodyStartHost; // worlds + 1
	DBuf<int32_t> bodyStart;
	std::vector<DiffStepRec*> steps;
	// reverse pass scratch
	DBuf<char> state;   // StepState arrays
	DBuf<dd::Node<float>> nodes;
	DBuf<float> adj;
	DBuf<Tape<float>> tapes;
	DBuf<uint32_t> outIds;
	DBuf<float> seed, leafAdj;
	DBuf<double> dstate, dparam;
	DBuf<float> fstate, fparam;
	DBuf<int32_t> scratchI;
	DBuf<uint32_t> counters;
	// parallel adjoint (k_adjoint.hip): per-world slabs, the recomputed
	// states of one checkpoint segment, seeds and gradients
	int checkpointEvery = 10;
	DBuf<float> adjSlab;
	DBuf<double> adjDSlab;
	DBuf<float> segStates;
	DBuf<double> seedA, seedB, massG, fricG;
	~DiffState()
	{
		adjSlab.release();
		adjDSlab.release();
		segStates.release();
		seedA.release();
		seedB.release();
		massG.release();
		fricG.release();
		for (DiffStepRec* r : steps)
		{
			r->release();
			delete r;
		}
		bodyStart.release();
		state.release();
		nodes.release();
		adj.release();
		tapes.release();
		outIds.release();
		seed.release();
		leafAdj.release();
		dstate.release();
		dparam.release();
		fstate.release();
		fparam.release();
		scratchI.release();
		counters.release();
	}
};

struct RecView
{
	const float4 *pos, *rot, *vel, *angVel, *invIL, *shape, *bmat;
	const float* angDamp;
	const uint32_t* flags;
	const int32_t* bodyStart;
	const int32_t* contactStart;
	const int2* bodies;
	const float4 *normal, *cmat, *pA, *pB, *pFA, *pFB, *pSlip;
	const int32_t* sched;
	int64_t cStride;
	int32_t worlds;
	float dt;
	int substeps;
	m3d_vec3 gravity;
	int iterations;
};

inline RecView rec_view(const DiffState* D, const DiffStepRec* R, int iterations)
{
	RecView V;
	V.pos = R->pos.p;
	V.rot = R->rot.p;
	V.vel = R->vel.p;
	V.angVel = R->angVel.p;
	V.invIL = R->invIL.p;
	V.shape = R->shape.p;
	V.bmat = R->bmat.p;
	V.angDamp = R->angDamp.p;
	V.flags = R->flags.p;
	V.bodyStart = D->bodyStart.p;
	V.contactStart = R->contactStart.p;
	V.bodies = R->bodies.p;
	V.normal = R->normal.p;
	V.cmat = R->cmat.p;
	V.pA = R->pA.p;
	V.pB = R->pB.p;
	V.pFA = R->pFA.p;
	V.pFB = R->pFB.p;
	V.pSlip = R->pSlip.p;
	V.sched = R->sched.p;
	V.cStride = R->nc;
	V.worlds = D->worlds;
	V.dt = R->dt;
	V.substeps = R->substeps;
	V.gravity = R->gravity;
	V.iterations = iterations;
	return V;
}

MX_HD StepIn world_in(const RecView& V, int32_t wi)
{
	int32_t b0 = V.bodyStart[wi], b1 = V.bodyStart[wi + 1];
	int32_t c0 = V.contactStart[wi], c1 = V.contactStart[wi + 1];
	StepIn in;
	in.nb = b1 - b0;
	in.pos = V.pos + b0;
	in.rot = V.rot + b0;
	in.vel = V.vel + b0;
	in.angVel = V.angVel + b0;
	in.invIL = V.invIL + b0;
	in.angDamp = V.angDamp + b0;
	in.shape = V.shape + b0;
	in.bmat = V.bmat + b0;
	in.flags = V.flags + b0;
	in.nc = c1 - c0;
	in.bodies = V.bodies + c0;
	in.normal = V.normal + c0;
	in.cmat = V.cmat + c0;
	in.pA = V.pA + 4 * (int64_t)c0;
	in.pB = V.pB + 4 * (int64_t)c0;
	in.pFA = V.pFA + 4 * (int64_t)c0;
	in.pFB = V.pFB + 4 * (int64_t)c0;
	in.pSlip = V.pSlip + 4 * (int64_t)c0;
	in.sched = V.sched != nullptr ? V.sched + c0 : nullptr;
	in.cStride = V.cStride;
	in.dt = V.dt;
	in.substeps = V.substeps;
	in.gravity = V.gravity;
	return in;
}

// replay scratch of one precision, carved from one byte buffer
template <class T>
struct StateBufs
{
	StepState<T> S;
	static size_t bytes(int64_t nb, int64_t nc)
	{
		int64_t np = 4 * nc;
		return (size_t)nb * (sizeof(dd::V3<T>) * 6 + sizeof(dd::Q4<T>) * 2 + sizeof(dd::M3<T>) + sizeof(T) +
							 sizeof(dd::V3<T>)) +
			   (size_t)nc * sizeof(T) + (size_t)np * (sizeof(T) * 3 + sizeof(dd::V3<T>) * 4 + sizeof(float4) * 3) +
			   64 * 32;
	}
	void carve(char* base, int64_t nb, int64_t nc)
	{
		int64_t np = 4 * nc;
		char* p = base;
		auto take = [&](size_t n) {
			char* r = p;
			p += (n + 63) & ~(size_t)63;
			return r;
		};
		S.x = (dd::V3<T>*)take(nb * sizeof(dd::V3<T>));
		S.q = (dd::Q4<T>*)take(nb * sizeof(dd::Q4<T>));
		S.v = (dd::V3<T>*)take(nb * sizeof(dd::V3<T>));
		S.w = (dd::V3<T>*)take(nb * sizeof(dd::V3<T>));
		S.px = (dd::V3<T>*)take(nb * sizeof(dd::V3<T>));
		S.pq = (dd::Q4<T>*)take(nb * sizeof(dd::Q4<T>));
		S.pv = (dd::V3<T>*)take(nb * sizeof(dd::V3<T>));
		S.pw = (dd::V3<T>*)take(nb * sizeof(dd::V3<T>));
		S.iw = (dd::M3<T>*)take(nb * sizeof(dd::M3<T>));
		S.invMass = (T*)take(nb * sizeof(T));
		S.il = (dd::V3<T>*)take(nb * sizeof(dd::V3<T>));
		S.mu = (T*)take(std::max<int64_t>(nc, 1) * sizeof(T));
		S.lambdaN = (T*)take(std::max<int64_t>(np, 1) * sizeof(T));
		S.lambdaT = (T*)take(std::max<int64_t>(np, 1) * sizeof(T));
		S.rA = (dd::V3<T>*)take(std::max<int64_t>(np, 1) * sizeof(dd::V3<T>));
		S.rB = (dd::V3<T>*)take(std::max<int64_t>(np, 1) * sizeof(dd::V3<T>));
		S.jA = (dd::V3<T>*)take(std::max<int64_t>(np, 1) * sizeof(dd::V3<T>));
		S.jB = (dd::V3<T>*)take(std::max<int64_t>(np, 1) * sizeof(dd::V3<T>));
		S.wN = (T*)take(std::max<int64_t>(np, 1) * sizeof(T));
		S.fA = (float4*)take(std::max<int64_t>(np, 1) * sizeof(float4));
		S.fB = (float4*)take(std::max<int64_t>(np, 1) * sizeof(float4));
		S.slip = (float4*)take(std::max<int64_t>(np, 1) * sizeof(float4));
	}
};

inline int iterations_of(const World& w) { return w.def.positionIterations > 0 ? w.def.positionIterations : 2; }

} // namespace mx
