// Adjoint of the XPBD solve over recorded steps (differentiable worlds).
//
// Recording: while enabled, every step stores, just before its solve, the
// body state and parameters and the solver contacts (ascending engine order,
// grouped by sub-world) with their manifolds: the fixed contact set of that
// step. The state recorded at each step start is the checkpoint the reverse
// pass starts from (a checkpoint per step costs 52 B per body).
//
// Reverse pass, step by step from the last: one lane per world replays the
// step's solve (mx_diff.h) on the reverse-mode tape, then walks the tape
// backwards from the adjoint of the step's output state, giving the adjoint
// of its input state (the seed of the step before) and the step's share of
// the mass and friction gradients.
#include "mx_diffrec.h"

#include <algorithm>
#include <string.h>
#include <vector>

namespace mx
{

DBodies make_body_view(World& w);
DContacts make_contact_view(World& w);

void diff_release(World& w)
{
	delete w.diff;
	w.diff = nullptr;
}

static inline unsigned gridk(int64_t n) { return (unsigned)((n + 255) / 256); }

// ------------------------------------------------------------ recording

__global__ void k_diff_flag(const uint32_t* contactList, int32_t NC, uint32_t* flag)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k < NC)
		flag[contactList[k]] = 1u;
}

__device__ __forceinline__ int32_t world_of_body(const int32_t* bodySub, int32_t b)
{
	return bodySub != nullptr ? bodySub[b] : 0;
}

__global__ void k_diff_world_counts(const uint32_t* flag, const int2* bodies, const float4* normal, int32_t C,
									const int32_t* bodySub, uint32_t* wc, uint32_t* wp)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c < C && flag[c])
	{
		int32_t sw = world_of_body(bodySub, bodies[c].x);
		atomicAdd(&wc[sw], 1u);
		atomicAdd(&wp[sw], 4u); // point slots: a substep's re-collision may fill all four
	}
}

__global__ void k_diff_gather(const int2* bodies, int32_t n, const uint32_t* flag, const uint32_t* offs,
							  const int32_t* bodySub, const int32_t* bodyStart, int2* ob, int32_t* idx)
{
	int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (c >= n || !flag[c])
		return;
	uint32_t k = offs[c];
	int2 ab = bodies[c];
	int32_t b0 = bodyStart[world_of_body(bodySub, ab.x)];
	ob[k] = make_int2(ab.x - b0, ab.y - b0);
	idx[k] = c;
}

// the manifold of every recorded contact after a substep's prepare, and its
// schedule word: the normal projection's form from the contact's solve
// record (bits 0-7: the packet's largest point count, or the serial form)
// and its pass-order key, the color of a colored island's contact (8: the
// overflow pass) or 0
__global__ void k_diff_gather_sub(DContacts C, const int32_t* idx, int32_t nc, int64_t base, const uint32_t* atOf,
								  const int4* rec, const int32_t* contactColor, float4* on, float4* om, float4* oA,
								  float4* oB, float4* oFA, float4* oFB, float4* oS, int32_t* osched)
{
	int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (k >= nc)
		return;
	int32_t c = idx[k];
	int64_t o = base + k;
	int32_t color = contactColor[c];
	osched[o] = (rec[atOf[c]].w & 0xFF) | ((color < 0 ? 0 : color) << 8);
	on[o] = C.normal[c];
	om[o] = C.mat[c];
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = pidx(C, c, j);
		oA[o * 4 + j] = C.pA[p];
		oB[o * 4 + j] = C.pB[p];
		oFA[o * 4 + j] = C.pFA[p];
		oFB[o * 4 + j] = C.pFB[p];
		oS[o * 4 + j] = C.pSlip[p];
	}
}

__global__ void k_diff_bodies(DBodies B, int32_t n, float* angDamp)
{
	int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n)
		angDamp[i] = B.iw0[i].w;
}

bool diff_record_step(World& w, const IslandLists& L, float dt, int substeps)
{
	DiffState* D = w.diff;
	if (D == nullptr || !D->recording)
		return true;
	// Joints are not replayed by the adjoint: such a step is not recorded,
	// recording stops, and the records refuse every later use, so no
	// gradient of a trajectory the engine did not run is ever returned. The
	// step itself runs as usual. Colored islands (src/xpbd.cpp:1167-1369) are
	// replayed with the colors and packet forms the solve used (recorded per
	// substep by diff_record_substep).
	if (w.djointCount > 0)
	{
		char msg[256];
		snprintf(msg, sizeof msg, "differentiable recording refused at step %d: the world has joints",
				 (int)D->steps.size());
		D->recording = false;
		D->refusedStep = (int32_t)D->steps.size();
		D->refusal = msg;
		set_error("%s", msg);
		return true;
	}
	hipStream_t s = w.stream;
	int32_t n = w.deviceBodyCount;
	int32_t C = w.contactCount;
	int32_t W = D->worlds;
	DiffStepRec* R = new DiffStepRec();
	D->steps.push_back(R);
	R->nb = n;
	R->dt = dt;
	R->substeps = substeps;
	R->gravity = w.def.gravity;
	DBodies B = make_body_view(w);
	if (!R->pos.ensure(n) || !R->rot.ensure(n) || !R->vel.ensure(n) || !R->angVel.ensure(n) || !R->invIL.ensure(n) ||
		!R->shape.ensure(n) || !R->bmat.ensure(n) || !R->angDamp.ensure(n) || !R->flags.ensure(n) ||
		!R->contactStart.ensure((size_t)W + 1))
		return false;
	hipMemcpyAsync(R->pos.p, B.pos, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
	hipMemcpyAsync(R->rot.p, B.rot, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
	hipMemcpyAsync(R->vel.p, B.vel, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
	hipMemcpyAsync(R->angVel.p, B.angVel, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
	hipMemcpyAsync(R->invIL.p, B.invIL, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
	hipMemcpyAsync(R->shape.p, B.shape, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
	hipMemcpyAsync(R->bmat.p, B.mat, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
	hipMemcpyAsync(R->flags.p, B.flags, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
	k_diff_bodies<<<gridk(n), 256, 0, s>>>(B, n, R->angDamp.p);
	// solver contacts in ascending order (contacts of a batch are grouped by
	// sub-world, so the compacted list is too)
	if (!D->counters.ensure(2 * (size_t)W + 2) || !w.u32a.ensure(std::max(C, 1)) || !w.u32b.ensure(std::max(C, 1)) ||
		!ensure_prim_scratch(w, std::max<int64_t>(C, W + 1)))
		return false;
	uint32_t* flag = w.u32a.p;
	uint32_t* offs = w.u32b.p;
	uint32_t* wc = D->counters.p;
	uint32_t* wp = wc + W + 1;
	launch_fill_u32(flag, 0u, std::max(C, 1), s);
	launch_fill_u32(wc, 0u, 2 * (int64_t)W + 2, s);
	if (L.NC > 0)
		k_diff_flag<<<gridk(L.NC), 256, 0, s>>>(L.contactList, L.NC, flag);
	int32_t NC = 0;
	if (C > 0)
	{
		k_diff_world_counts<<<gridk(C), 256, 0, s>>>(flag, w.cc.bodies.p, w.cc.normal.p, C,
													 w.subCount > 0 ? w.bodySub.p : nullptr, wc, wp);
		scan_exclusive(flag, offs, C, w.counters.p + CT_DIFF_CONTACTS, w.scanStatus.p, s);
	}
	scan_exclusive(wc, (uint32_t*)R->contactStart.p, W + 1, nullptr, w.scanStatus.p, s);
	std::vector<uint32_t> wpHost((size_t)W);
	hipMemcpyAsync(wpHost.data(), wp, W * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
	if (!sync_counters(w))
		return false;
	NC = C > 0 ? (int32_t)w.hostCounters[CT_DIFF_CONTACTS] : 0;
	R->nc = NC;
	for (int32_t k = 0; k < W; ++k)
	{
		R->maxWorldPoints = std::max<int32_t>(R->maxWorldPoints, (int32_t)wpHost[k]);
		R->maxWorldBodies = std::max<int32_t>(R->maxWorldBodies, D->bodyStartHost[k + 1] - D->bodyStartHost[k]);
	}
	size_t nc1 = std::max(NC, 1);
	size_t ncs = nc1 * (size_t)substeps;
	if (!R->bodies.ensure(nc1) || !R->idx.ensure(nc1) || !R->normal.ensure(ncs) || !R->cmat.ensure(ncs) ||
		!R->pA.ensure(4 * ncs) || !R->pB.ensure(4 * ncs) || !R->pFA.ensure(4 * ncs) || !R->pFB.ensure(4 * ncs) ||
		!R->pSlip.ensure(4 * ncs) || !R->sched.ensure(ncs))
		return false;
	if (NC > 0)
		k_diff_gather<<<gridk(C), 256, 0, s>>>(w.cc.bodies.p, C, flag, offs, w.subCount > 0 ? w.bodySub.p : nullptr,
											   D->bodyStart.p, R->bodies.p, R->idx.p);
	return hip_ok(hipStreamSynchronize(s), "step record");
}

bool diff_is_recording(const World& w) { return w.diff != nullptr && w.diff->recording; }

bool diff_refused(const World& w)
{
	if (w.diff == nullptr || w.diff->refusedStep < 0)
		return false;
	set_error("%s", w.diff->refusal.c_str());
	return true;
}

bool diff_record_substep(World& w, int sub, const uint32_t* atOf, const int4* rec)
{
	DiffState* D = w.diff;
	if (D == nullptr || !D->recording || D->steps.empty())
		return true;
	DiffStepRec* R = D->steps.back();
	if (R->nc == 0 || sub >= R->substeps)
		return true;
	int64_t base = (int64_t)sub * R->nc;
	k_diff_gather_sub<<<gridk(R->nc), 256, 0, w.stream>>>(make_contact_view(w), R->idx.p, R->nc, base, atOf, rec,
														  w.contactColor.p, R->normal.p, R->cmat.p, R->pA.p, R->pB.p,
														  R->pFA.p, R->pFB.p, R->pSlip.p, R->sched.p);
	return true;
}

// ------------------------------------------------------------ replay

template <class T>
__device__ StepState<T> world_state(const StepState<T>& G, int32_t b0, int32_t c0)
{
	StepState<T> S = G;
	S.x += b0;
	S.q += b0;
	S.v += b0;
	S.w += b0;
	S.px += b0;
	S.pq += b0;
	S.pv += b0;
	S.pw += b0;
	S.iw += b0;
	S.invMass += b0;
	S.il += b0;
	S.mu += c0;
	int64_t p0 = 4 * (int64_t)c0;
	S.lambdaN += p0;
	S.lambdaT += p0;
	S.rA += p0;
	S.rB += p0;
	S.jA += p0;
	S.jB += p0;
	S.wN += p0;
	S.fA += p0;
	S.fB += p0;
	S.slip += p0;
	return S;
}

// contact friction from the two bodies' friction (src/world.cpp:1554,
// k_create: sqrt of the product, lower body index first)
template <class T>
__device__ void init_contacts(const StepIn& in, StepState<T>& S, const T* bodyMu)
{
	for (int32_t c = 0; c < in.nc; ++c)
	{
		int2 ab = in.bodies[c];
		S.mu[c] = dd::sqrt_(bodyMu[ab.x] * bodyMu[ab.y]);
	}
}

template <class T>
__device__ void init_inertia(const StepIn& in, StepState<T>& S, int32_t i)
{
	S.iw[i] = dd::inv_inertia_world(S.q[i], S.il[i]);
	S.px[i] = S.x[i];
	S.pq[i] = S.q[i];
	S.pv[i] = S.v[i];
	S.pw[i] = S.w[i];
}

// Plain-precision replay of one step for every world (one lane per world).
// state: per body 13 values (x, q, v, w) in, then out; null in = recorded.
// mass, friction: per body, null = recorded.
template <class T>
__global__ void __launch_bounds__(64) k_diff_replay(RecView V, StepState<T> G, const T* stateIn, T* stateOut, const T* mass, const T* mu,
							  T* muScratch)
{
	// one world per wave: worlds branch differently, so they do not share one
	int32_t wi = (int32_t)blockIdx.x;
	if (wi >= V.worlds || threadIdx.x != 0)
		return;
	StepIn in = world_in(V, wi);
	int32_t b0 = V.bodyStart[wi];
	StepState<T> S = world_state(G, b0, V.contactStart[wi]);
	T* bmu = muScratch + b0;
	for (int32_t i = 0; i < in.nb; ++i)
	{
		int32_t g = b0 + i;
		float4 P = in.pos[i], Q = in.rot[i], Vv = in.vel[i], Ww = in.angVel[i], IL = in.invIL[i];
		if (stateIn != nullptr)
		{
			const T* s = stateIn + (int64_t)g * kOuts;
			S.x[i] = dd::v3<T>(s[0], s[1], s[2]);
			S.q[i] = { s[3], s[4], s[5], s[6] };
			S.v[i] = dd::v3<T>(s[7], s[8], s[9]);
			S.w[i] = dd::v3<T>(s[10], s[11], s[12]);
		}
		else
		{
			S.x[i] = dd::v3<T>(P.x, P.y, P.z);
			S.q[i] = { (T)Q.x, (T)Q.y, (T)Q.z, (T)Q.w };
			S.v[i] = dd::v3<T>(Vv.x, Vv.y, Vv.z);
			S.w[i] = dd::v3<T>(Ww.x, Ww.y, Ww.z);
		}
		T invMass = (T)P.w;
		dd::V3<T> il = dd::v3<T>(IL.x, IL.y, IL.z);
		if (mass != nullptr && P.w != 0.0f)
		{
			// inverse mass and inertia scale as 1 / mass
			T m = mass[g];
			T k = (T)Ww.w / m;
			invMass = (T)1 / m;
			il = dd::v3<T>(il.x * k, il.y * k, il.z * k);
		}
		S.invMass[i] = invMass;
		S.il[i] = il;
		bmu[i] = mu != nullptr ? mu[g] : (T)in.bmat[i].y;
		init_inertia(in, S, i);
	}
	init_contacts(in, S, bmu);
	dd::solve_step(in, S, V.iterations);
	for (int32_t i = 0; i < in.nb; ++i)
	{
		T* o = stateOut + (int64_t)(b0 + i) * kOuts;
		o[0] = S.x[i].x;
		o[1] = S.x[i].y;
		o[2] = S.x[i].z;
		o[3] = S.q[i].x;
		o[4] = S.q[i].y;
		o[5] = S.q[i].z;
		o[6] = S.q[i].w;
		o[7] = S.v[i].x;
		o[8] = S.v[i].y;
		o[9] = S.v[i].z;
		o[10] = S.w[i].x;
		o[11] = S.w[i].y;
		o[12] = S.w[i].z;
	}
}

// Taped replay of one step (float values), from the recorded state: leaves
// per body x, q, v, w, mass, friction (ids i * 15 + k, in that order), the
// output node of every state component in outIds.
__global__ void __launch_bounds__(64) k_diff_tape(RecView V, StepState<Ad<float>> G, Tape<float>* tapes, uint32_t* outIds,
							Ad<float>* muScratch)
{
	typedef Ad<float> A;
	int32_t wi = (int32_t)blockIdx.x;
	if (wi >= V.worlds || threadIdx.x != 0)
		return;
	Tape<float>& tl = dd::ad_tape(); // this lane's tape header, in LDS
	tl = tapes[wi];
	StepIn in = world_in(V, wi);
	int32_t b0 = V.bodyStart[wi];
	StepState<A> S = world_state(G, b0, V.contactStart[wi]);
	A* bmu = muScratch + b0;
	// leaves first: body i's k-th leaf is node i * 15 + k
	for (int32_t i = 0; i < in.nb; ++i)
	{
		float4 P = in.pos[i], Q = in.rot[i], Vv = in.vel[i], Ww = in.angVel[i];
		A x = dd::ad_leaf(P.x), y = dd::ad_leaf(P.y), z = dd::ad_leaf(P.z);
		A qx = dd::ad_leaf(Q.x), qy = dd::ad_leaf(Q.y), qz = dd::ad_leaf(Q.z), qw = dd::ad_leaf(Q.w);
		A vx = dd::ad_leaf(Vv.x), vy = dd::ad_leaf(Vv.y), vz = dd::ad_leaf(Vv.z);
		A wx = dd::ad_leaf(Ww.x), wy = dd::ad_leaf(Ww.y), wz = dd::ad_leaf(Ww.z);
		S.x[i] = dd::v3<A>(x, y, z);
		S.q[i] = { qx, qy, qz, qw };
		S.v[i] = dd::v3<A>(vx, vy, vz);
		S.w[i] = dd::v3<A>(wx, wy, wz);
		S.invMass[i] = dd::ad_leaf(Ww.w); // the mass leaf, turned into 1 / m below
		bmu[i] = dd::ad_leaf(in.bmat[i].y);
	}
	for (int32_t i = 0; i < in.nb; ++i)
	{
		float4 P = in.pos[i], Ww = in.angVel[i], IL = in.invIL[i];
		A m = S.invMass[i];
		// the recorded inverse mass and inertia, with d/dm = -value / m
		float mv = Ww.w;
		if (P.w != 0.0f && mv > 0.0f)
		{
			S.invMass[i] = dd::ad_make<float>(P.w, m.id, -P.w / mv, dd::kNoNode, 0.0f);
			S.il[i] = dd::v3<A>(dd::ad_make<float>(IL.x, m.id, -IL.x / mv, dd::kNoNode, 0.0f),
								dd::ad_make<float>(IL.y, m.id, -IL.y / mv, dd::kNoNode, 0.0f),
								dd::ad_make<float>(IL.z, m.id, -IL.z / mv, dd::kNoNode, 0.0f));
		}
		else
		{
			S.invMass[i] = dd::ad_const(P.w);
			S.il[i] = dd::v3<A>(dd::ad_const(IL.x), dd::ad_const(IL.y), dd::ad_const(IL.z));
		}
		init_inertia(in, S, i);
	}
	init_contacts(in, S, bmu);
	dd::solve_step(in, S, V.iterations);
	for (int32_t i = 0; i < in.nb; ++i)
	{
		uint32_t* o = outIds + (int64_t)(b0 + i) * kOuts;
		o[0] = S.x[i].x.id;
		o[1] = S.x[i].y.id;
		o[2] = S.x[i].z.id;
		o[3] = S.q[i].x.id;
		o[4] = S.q[i].y.id;
		o[5] = S.q[i].z.id;
		o[6] = S.q[i].w.id;
		o[7] = S.v[i].x.id;
		o[8] = S.v[i].y.id;
		o[9] = S.v[i].z.id;
		o[10] = S.w[i].x.id;
		o[11] = S.w[i].y.id;
		o[12] = S.w[i].z.id;
	}
	tapes[wi].n = tl.n;
	tapes[wi].overflow = tl.overflow;
}

// Reverse sweep of each world's tape: adj(out) = seed, then every node
// passes its adjoint to its inputs; the leaves' adjoints are the gradient.
// Inputs are mostly recent nodes, so the adjoints of the kWin nodes below
// the current one live in an LDS ring (slot k % kWin) and only older
// inputs go to the global array.
constexpr uint32_t kWin = 8192;

__global__ void k_diff_backprop(const Tape<float>* tapes, const int32_t* bodyStart, int32_t worlds, float* adj,
								const uint32_t* outIds, const float* seed, float* leafAdj)
{
	__shared__ float win[kWin];
	int32_t wi = (int32_t)blockIdx.x;
	if (wi >= worlds)
		return;
	for (uint32_t q = threadIdx.x; q < kWin; q += blockDim.x)
		win[q] = 0.0f;
	__syncthreads();
	if (threadIdx.x != 0)
		return;
	const Tape<float>& T = tapes[wi];
	const int32_t b0 = bodyStart[wi], b1 = bodyStart[wi + 1];
	const uint32_t n = (uint32_t)T.n;
	const dd::Node<float>* nodes = T.nodes;
	float* glob = adj + (T.nodes - tapes[0].nodes); // this world's part
	// nodes below the first window get contributions in glob
	const uint32_t low = n > kWin ? n - kWin : 0u;
	for (uint32_t k = 0; k < low; ++k)
		glob[k] = 0.0f;
	for (int32_t g = b0; g < b1; ++g)
		for (int k = 0; k < kOuts; ++k)
		{
			uint32_t id = outIds[(int64_t)g * kOuts + k];
			if (id == dd::kNoNode || id >= n)
				continue;
			float sd = seed[(int64_t)g * kOuts + k];
			if (id >= low)
				win[id % kWin] += sd;
			else
				glob[id] += sd;
		}
	const uint32_t nLeaves = (uint32_t)(b1 - b0) * kLeaves;
	for (uint32_t k = n; k-- > 0;)
	{
		// node k's adjoint: its ring slot, plus what reached it while it was
		// below the ring (only nodes below `low` ever were)
		float a = win[k % kWin];
		win[k % kWin] = 0.0f; // from here on the slot is node k - kWin's
		if (k < low)
			a += glob[k];
		if (k < nLeaves)
			leafAdj[(int64_t)(b0 + k / kLeaves) * kLeaves + k % kLeaves] = a;
		if (a == 0.0f)
			continue;
		dd::Node<float> nd = nodes[k];
		// after node k the ring holds nodes [k - kWin, k - 1]
		uint32_t edge = k >= kWin ? k - kWin : 0u;
		if (nd.a != dd::kNoNode)
		{
			if (nd.a >= edge)
				win[nd.a % kWin] += a * nd.da;
			else
				glob[nd.a] += a * nd.da;
		}
		if (nd.b != dd::kNoNode)
		{
			if (nd.b >= edge)
				win[nd.b % kWin] += a * nd.db;
			else
				glob[nd.b] += a * nd.db;
		}
	}
}

__global__ void k_diff_tapes_init(Tape<float>* tapes, int32_t worlds, dd::Node<float>* nodes, uint32_t cap)
{
	int32_t wi = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
	if (wi >= worlds)
		return;
	// each world's nodes contiguous: the reverse sweep reads them in order
	Tape<float> t;
	t.nodes = nodes + (uint64_t)wi * cap;
	t.stride = 1;
	t.n = 0;
	t.cap = cap;
	t.overflow = 0;
	tapes[wi] = t;
}

// ------------------------------------------------------------ host


template <class T>
static bool ensure_state(DiffState* D, const DiffStepRec* R, StepState<T>* out)
{
	StateBufs<T> sb;
	size_t bytes = StateBufs<T>::bytes(R->nb, std::max(R->nc, 1));
	if (!D->state.ensure(bytes))
		return false;
	sb.carve(D->state.p, R->nb, std::max(R->nc, 1));
	*out = sb.S;
	return true;
}


// Forward tangents of the recorded steps on the host (see m3dx_diff_jvp_cpu):
// values in P (float: the engine's own path, double: the FP64 replay),
// derivatives in double.
template <class P>
static int jvp_host(World& w, int first, int count, const double* stateIn, const double* dStateIn, const double* mass,
					const double* dMass, const double* friction, const double* dFriction, double* stateOut,
					double* dStateOut)
{
	typedef dd::Dual<P, double> D;
	DiffState* Ds = w.diff;
	const int32_t n = Ds->steps[first]->nb;
	const int32_t W = Ds->worlds;
	const int iterations = iterations_of(w);
	std::vector<D> cur((size_t)n * kOuts);
	for (size_t k = 0; k < cur.size(); ++k)
		cur[k] = dd::dual<P, double>((P)stateIn[k], dStateIn[k]);
	auto down = [](auto& v, const auto& buf, size_t cnt) {
		v.resize(cnt);
		return cnt == 0 || hipMemcpy(v.data(), buf.p, cnt * sizeof(v[0]), hipMemcpyDeviceToHost) == hipSuccess;
	};
	std::vector<float4> pos, rot, vel, angVel, invIL, shape, bmat, normal, cmat, pA, pB, pFA, pFB, pSlip;
	std::vector<float> angDamp;
	std::vector<uint32_t> flags;
	std::vector<int32_t> cstart;
	std::vector<int2> bodies;
	std::vector<int32_t> sched;
	std::vector<char> mem;
	for (int k = 0; k < count; ++k)
	{
		const DiffStepRec* R = Ds->steps[first + k];
		if (R->nb != n)
		{
			set_error("body count changed while recording");
			return -1;
		}
		size_t nb = (size_t)n, ncs = (size_t)R->nc * R->substeps;
		bool ok = down(pos, R->pos, nb) && down(rot, R->rot, nb) && down(vel, R->vel, nb) && down(angVel, R->angVel, nb) &&
				  down(invIL, R->invIL, nb) && down(shape, R->shape, nb) && down(bmat, R->bmat, nb) &&
				  down(angDamp, R->angDamp, nb) && down(flags, R->flags, nb) &&
				  down(cstart, R->contactStart, (size_t)W + 1) && down(bodies, R->bodies, (size_t)R->nc) &&
				  down(normal, R->normal, ncs) && down(cmat, R->cmat, ncs) && down(pA, R->pA, 4 * ncs) &&
				  down(pB, R->pB, 4 * ncs) && down(pFA, R->pFA, 4 * ncs) && down(pFB, R->pFB, 4 * ncs) &&
				  down(pSlip, R->pSlip, 4 * ncs) && down(sched, R->sched, ncs);
		if (!ok)
		{
			set_error("jvp: record download failed");
			return -1;
		}
		for (int32_t wi = 0; wi < W; ++wi)
		{
			int32_t b0 = Ds->bodyStartHost[wi], b1 = Ds->bodyStartHost[wi + 1];
			int32_t c0 = cstart[wi], c1 = cstart[wi + 1];
			StepIn in;
			in.nb = b1 - b0;
			in.pos = pos.data() + b0;
			in.rot = rot.data() + b0;
			in.vel = vel.data() + b0;
			in.angVel = angVel.data() + b0;
			in.invIL = invIL.data() + b0;
			in.angDamp = angDamp.data() + b0;
			in.shape = shape.data() + b0;
			in.bmat = bmat.data() + b0;
			in.flags = flags.data() + b0;
			in.nc = c1 - c0;
			in.bodies = bodies.data() + c0;
			in.normal = normal.data() + c0;
			in.cmat = cmat.data() + c0;
			in.pA = pA.data() + 4 * (int64_t)c0;
			in.pB = pB.data() + 4 * (int64_t)c0;
			in.pFA = pFA.data() + 4 * (int64_t)c0;
			in.pFB = pFB.data() + 4 * (int64_t)c0;
			in.pSlip = pSlip.data() + 4 * (int64_t)c0;
			in.sched = sched.data() + c0;
			in.cStride = R->nc;
			in.dt = R->dt;
			in.substeps = R->substeps;
			in.gravity = R->gravity;
			StateBufs<D> sb;
			mem.assign(StateBufs<D>::bytes(in.nb, std::max(in.nc, 1)), 0);
			sb.carve(mem.data(), in.nb, std::max(in.nc, 1));
			StepState<D>& S = sb.S;
			std::vector<D> bmu((size_t)in.nb);
			for (int32_t i = 0; i < in.nb; ++i)
			{
				const D* x = cur.data() + (size_t)(b0 + i) * kOuts;
				S.x[i] = dd::v3<D>(x[0], x[1], x[2]);
				S.q[i] = { x[3], x[4], x[5], x[6] };
				S.v[i] = dd::v3<D>(x[7], x[8], x[9]);
				S.w[i] = dd::v3<D>(x[10], x[11], x[12]);
				float4 Pp = in.pos[i], IL = in.invIL[i];
				D invMass = dd::dual<P, double>((P)Pp.w, 0.0);
				dd::V3<D> il = dd::v3<D>(dd::dual<P, double>((P)IL.x, 0.0), dd::dual<P, double>((P)IL.y, 0.0),
										 dd::dual<P, double>((P)IL.z, 0.0));
				if (Pp.w != 0.0f && mass != nullptr)
				{
					// inverse mass and inertia scale as 1 / mass
					D m = dd::dual<P, double>((P)mass[b0 + i], dMass[b0 + i]);
					D kk = (P)in.angVel[i].w / m;
					invMass = (P)1 / m;
					il = dd::v3<D>(il.x * kk, il.y * kk, il.z * kk);
				}
				else if (Pp.w != 0.0f && in.angVel[i].w > 0.0f)
				{
					// the recorded values, with d/dm = -value / m
					double m = in.angVel[i].w, dm = dMass[b0 + i];
					invMass.d = -(double)invMass.v / m * dm;
					il.x.d = -(double)il.x.v / m * dm;
					il.y.d = -(double)il.y.v / m * dm;
					il.z.d = -(double)il.z.v / m * dm;
				}
				S.invMass[i] = invMass;
				S.il[i] = il;
				bmu[i] = dd::dual<P, double>(friction != nullptr ? (P)friction[b0 + i] : (P)in.bmat[i].y,
											 dFriction[b0 + i]);
				S.iw[i] = dd::inv_inertia_world(S.q[i], S.il[i]);
				S.px[i] = S.x[i];
				S.pq[i] = S.q[i];
				S.pv[i] = S.v[i];
				S.pw[i] = S.w[i];
			}
			for (int32_t c = 0; c < in.nc; ++c)
			{
				int2 ab = in.bodies[c];
				S.mu[c] = dd::sqrt_(bmu[ab.x] * bmu[ab.y]);
			}
			dd::solve_step(in, S, iterations);
			for (int32_t i = 0; i < in.nb; ++i)
			{
				D* o = cur.data() + (size_t)(b0 + i) * kOuts;
				D vals[kOuts] = { S.x[i].x, S.x[i].y, S.x[i].z, S.q[i].x, S.q[i].y, S.q[i].z, S.q[i].w,
								  S.v[i].x, S.v[i].y, S.v[i].z, S.w[i].x, S.w[i].y, S.w[i].z };
				for (int q = 0; q < kOuts; ++q)
					o[q] = vals[q];
			}
		}
	}
	for (size_t k = 0; k < cur.size(); ++k)
	{
		stateOut[k] = (double)cur[k].v;
		dStateOut[k] = cur[k].d;
	}
	return 0;
}

} // namespace mx

// ------------------------------------------------------------ C ABI

using namespace mx;

extern "C"
{

// Starts (enable != 0: previous records dropped) or stops recording steps
// for the adjoint. Recording needs sleep disabled (a step's output must be
// its solve's output). Returns 0 or -1.
int m3dx_diff_record(m3d_world* world, int enable)
{
	World& w = world->w;
	if (enable)
	{
		if (w.def.enableSleep)
		{
			set_error("differentiable recording needs a world created with enableSleep = false");
			return -1;
		}
		diff_release(w);
		w.diff = new DiffState();
		DiffState* D = w.diff;
		D->worlds = w.subCount > 0 ? w.subCount : 1;
		if (w.subCount > 0)
			D->bodyStartHost = w.subBodyStart;
		else
			D->bodyStartHost = { 0, (int32_t)w.bodies.size() };
		if (!D->bodyStart.ensure(D->bodyStartHost.size()) ||
			!hip_ok(hipMemcpy(D->bodyStart.p, D->bodyStartHost.data(), D->bodyStartHost.size() * sizeof(int32_t),
							  hipMemcpyHostToDevice),
					"record setup"))
			return -1;
		D->recording = true;
		return 0;
	}
	if (w.diff != nullptr)
		w.diff->recording = false;
	return diff_refused(w) ? -1 : 0;
}

int32_t m3dx_diff_steps(m3d_world* world)
{
	World& w = world->w;
	return w.diff != nullptr ? (int32_t)w.diff->steps.size() : 0;
}

// The state recorded at the start of a step (13 floats per body: x q v w)
// and its solver contact count. Returns the contact count or -1.
int m3dx_diff_state(m3d_world* world, int step, float* out)
{
	World& w = world->w;
	if (diff_refused(w))
		return -1;
	DiffState* D = w.diff;
	if (D == nullptr || step < 0 || step >= (int)D->steps.size())
		return -1;
	const DiffStepRec* R = D->steps[step];
	std::vector<float4> p(R->nb), q(R->nb), v(R->nb), a(R->nb);
	hipMemcpy(p.data(), R->pos.p, R->nb * sizeof(float4), hipMemcpyDeviceToHost);
	hipMemcpy(q.data(), R->rot.p, R->nb * sizeof(float4), hipMemcpyDeviceToHost);
	hipMemcpy(v.data(), R->vel.p, R->nb * sizeof(float4), hipMemcpyDeviceToHost);
	hipMemcpy(a.data(), R->angVel.p, R->nb * sizeof(float4), hipMemcpyDeviceToHost);
	for (int32_t i = 0; i < R->nb; ++i)
	{
		float* o = out + (size_t)i * kOuts;
		o[0] = p[i].x;
		o[1] = p[i].y;
		o[2] = p[i].z;
		o[3] = q[i].x;
		o[4] = q[i].y;
		o[5] = q[i].z;
		o[6] = q[i].w;
		o[7] = v[i].x;
		o[8] = v[i].y;
		o[9] = v[i].z;
		o[10] = a[i].x;
		o[11] = a[i].y;
		o[12] = a[i].z;
	}
	return R->nc;
}

// Replays the recorded steps first .. first + count - 1 in sequence (one
// lane per world), starting from the state recorded at step `first` (or
// from stateIn), with per-body mass and friction (null: as recorded).
// precision 0: float (the engine's own arithmetic), 1: double. Arrays are
// host memory of the precision's type: stateIn/stateOut 13 per body (x, q,
// v, w), mass and friction 1 per body. Returns 0 or -1.
int m3dx_diff_replay(m3d_world* world, int first, int count, int precision, const void* stateIn, const void* mass,
					 const void* friction, void* stateOut)
{
	World& w = world->w;
	if (diff_refused(w))
		return -1;
	DiffState* D = w.diff;
	if (D == nullptr || first < 0 || count < 1 || first + count > (int)D->steps.size())
	{
		set_error("no recorded steps %d..%d", first, first + count - 1);
		return -1;
	}
	hipStream_t s = w.stream;
	int32_t n = D->steps[first]->nb;
	size_t el = precision ? sizeof(double) : sizeof(float);
	size_t stateBytes = (size_t)n * kOuts * el;
	DBuf<char> buf;
	if (!buf.ensure(2 * stateBytes + 3 * (size_t)n * el + 256))
		return -1;
	char* sA = buf.p;
	char* sB = sA + stateBytes;
	char* pm = sB + stateBytes;
	char* pf = pm + (size_t)n * el;
	char* bmu = pf + (size_t)n * el;
	if (stateIn)
		hipMemcpy(sA, stateIn, stateBytes, hipMemcpyHostToDevice);
	if (mass)
		hipMemcpy(pm, mass, (size_t)n * el, hipMemcpyHostToDevice);
	if (friction)
		hipMemcpy(pf, friction, (size_t)n * el, hipMemcpyHostToDevice);
	bool ok = true;
	for (int k = 0; k < count && ok; ++k)
	{
		const DiffStepRec* R = D->steps[first + k];
		RecView V = rec_view(D, R, iterations_of(w));
		bool useIn = k > 0 || stateIn != nullptr;
		if (precision)
		{
			StepState<double> G;
			ok = ensure_state(D, R, &G);
			if (ok)
				k_diff_replay<double><<<D->worlds, 64, 0, s>>>(
					V, G, useIn ? (const double*)sA : nullptr, (double*)sB, mass ? (const double*)pm : nullptr,
					friction ? (const double*)pf : nullptr, (double*)bmu);
		}
		else
		{
			StepState<float> G;
			ok = ensure_state(D, R, &G);
			if (ok)
				k_diff_replay<float><<<D->worlds, 64, 0, s>>>(
					V, G, useIn ? (const float*)sA : nullptr, (float*)sB, mass ? (const float*)pm : nullptr,
					friction ? (const float*)pf : nullptr, (float*)bmu);
		}
		ok = ok && hip_ok(hipStreamSynchronize(s), "replay");
		std::swap(sA, sB);
	}
	if (ok)
		ok = hip_ok(hipMemcpy(stateOut, sA, stateBytes, hipMemcpyDeviceToHost), "replay output");
	buf.release();
	return ok ? 0 : -1;
}

// Vector-Jacobian product of the recorded steps: given the adjoint of the
// final state (host, 13 per body: x, q, v, w), returns the adjoint of the
// first recorded step's input state (13 per body; its v is the gradient with
// respect to the initial velocities) and the gradients with respect to each
// body's mass and friction (summed over the steps). Any pointer but seed may
// be null. Returns 0, or -1 (m3dx_last_error) on tape overflow or error.
int m3dx_diff_backward_tape(m3d_world* world, const float* seed, float* stateAdj, float* massGrad, float* frictionGrad)
{
	World& w = world->w;
	if (diff_refused(w))
		return -1;
	DiffState* D = w.diff;
	if (D == nullptr || D->steps.empty())
	{
		set_error("no recorded steps");
		return -1;
	}
	hipStream_t s = w.stream;
	int32_t W = D->worlds;
	int32_t n = D->steps[0]->nb;
	if (!D->seed.ensure((size_t)n * kOuts) || !D->leafAdj.ensure((size_t)n * kLeaves) ||
		!D->outIds.ensure((size_t)n * kOuts) || !D->tapes.ensure(W))
		return -1;
	hipMemcpy(D->seed.p, seed, (size_t)n * kOuts * sizeof(float), hipMemcpyHostToDevice);
	std::vector<float> mg((size_t)n, 0.0f), fg((size_t)n, 0.0f), la((size_t)n * kLeaves);
	std::vector<Tape<float>> th((size_t)W);
	for (int step = (int)D->steps.size() - 1; step >= 0; --step)
	{
		const DiffStepRec* R = D->steps[step];
		if (R->nb != n)
		{
			set_error("body count changed while recording");
			return -1;
		}
		// tape capacity per world from the step's largest world
		uint64_t cap = (uint64_t)R->maxWorldBodies * (kLeaves + 1000u) + (uint64_t)R->maxWorldPoints * 4500u + 4096u;
		if (cap > 0xFFFFFFF0ull)
			cap = 0xFFFFFFF0ull;
		size_t entries = (size_t)cap * W;
		if (!D->nodes.ensure(entries) || !D->adj.ensure(entries))
			return -1;
		k_diff_tapes_init<<<gridk(W), 256, 0, s>>>(D->tapes.p, W, D->nodes.p, (uint32_t)cap);
		StepState<Ad<float>> G;
		if (!ensure_state(D, R, &G) || !D->scratchI.ensure((size_t)n * sizeof(Ad<float>) / sizeof(int32_t) + 4))
			return -1;
		RecView V = rec_view(D, R, iterations_of(w));
		k_diff_tape<<<W, 64, 0, s>>>(V, G, D->tapes.p, D->outIds.p, (Ad<float>*)D->scratchI.p);
		k_diff_backprop<<<W, 64, 0, s>>>(D->tapes.p, D->bodyStart.p, W, D->adj.p, D->outIds.p, D->seed.p,
												D->leafAdj.p);
		hipMemcpyAsync(th.data(), D->tapes.p, W * sizeof(Tape<float>), hipMemcpyDeviceToHost, s);
		hipMemcpyAsync(la.data(), D->leafAdj.p, la.size() * sizeof(float), hipMemcpyDeviceToHost, s);
		if (!hip_ok(hipStreamSynchronize(s), "adjoint step"))
			return -1;
		for (const Tape<float>& t : th)
			if (t.overflow)
			{
				set_error("adjoint tape overflow at step %d (capacity %llu nodes per world)", step,
						  (unsigned long long)cap);
				return -1;
			}
		// next seed: the adjoint of this step's input state
		std::vector<float> ns((size_t)n * kOuts);
		for (int32_t i = 0; i < n; ++i)
		{
			for (int k = 0; k < kOuts; ++k)
				ns[(size_t)i * kOuts + k] = la[(size_t)i * kLeaves + k];
			mg[i] += la[(size_t)i * kLeaves + 13];
			fg[i] += la[(size_t)i * kLeaves + 14];
		}
		hipMemcpy(D->seed.p, ns.data(), ns.size() * sizeof(float), hipMemcpyHostToDevice);
		if (step == 0 && stateAdj != nullptr)
			memcpy(stateAdj, ns.data(), ns.size() * sizeof(float));
	}
	if (massGrad != nullptr)
		memcpy(massGrad, mg.data(), mg.size() * sizeof(float));
	if (frictionGrad != nullptr)
		memcpy(frictionGrad, fg.data(), fg.size() * sizeof(float));
	return 0;
}

// Forward tangents of the recorded steps first .. first + count - 1 on the
// host CPU (the same solve templates as the device replays) from stateIn with
// per-body mass and friction (null: the recorded values), along one
// direction (dStateIn, dMass, dFriction); derivatives in double. primal 0:
// values in float, the engine's own path (the float steps' derivative in
// double precision); 1: values in double (the FP64 replay, whose path may
// leave the engine's where a branch is close). stateOut and dStateOut
// receive the final state and its directional derivative (13 per body). Host
// arrays of doubles; returns 0 or -1.
int m3dx_diff_jvp_cpu(m3d_world* world, int first, int count, int primal, const double* stateIn,
					  const double* dStateIn, const double* mass, const double* dMass, const double* friction,
					  const double* dFriction, double* stateOut, double* dStateOut)
{
	World& w = world->w;
	if (diff_refused(w))
		return -1;
	DiffState* Ds = w.diff;
	if (Ds == nullptr || first < 0 || count < 1 || first + count > (int)Ds->steps.size() || stateIn == nullptr ||
		dStateIn == nullptr || dMass == nullptr || dFriction == nullptr || stateOut == nullptr || dStateOut == nullptr)
	{
		set_error("jvp: no recorded steps %d..%d or a null array", first, first + count - 1);
		return -1;
	}
	if (primal)
		return jvp_host<double>(w, first, count, stateIn, dStateIn, mass, dMass, friction, dFriction, stateOut,
								dStateOut);
	return jvp_host<float>(w, first, count, stateIn, dStateIn, mass, dMass, friction, dFriction, stateOut, dStateOut);
}

} // extern "C"
