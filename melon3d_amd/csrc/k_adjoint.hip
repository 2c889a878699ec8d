// Parallel adjoint of the recorded XPBD solve (differentiable worlds).
//
// One workgroup per world and recorded step. The workgroup first replays the
// step's solve forward in float with the ops of mx_diff.h, which is the
// engine's arithmetic in the engine's order: the contact order of a pass
// (serial islands: src/xpbd.cpp:898-974; colored islands: their colors in
// turn, then the overflow contacts, :1269-1368, the colored normal
// projection in its packet form) is split into levels, a contact's level
// being one more than the highest level of the earlier contacts that share
// one of its dynamic bodies, so the contacts of one level write disjoint
// bodies and run in parallel with the serial result. Every op's inputs are
// recorded.
//
// The substeps are then walked backwards, each pass's levels in reverse. The
// vector-Jacobian product of one op (one contact or body) is assembled from
// forward tangents: one lane per (op, input component) evaluates the op on
// Dual<float> with that component's tangent set to one and dots the output
// tangents with the outputs' adjoints, which gives that component's adjoint.
// All lanes of a level read before any writes (a workgroup barrier between
// the two phases); a state component the op writes takes the product
// (replace), one it only reads or one that is read-only in the pass
// (inverse mass and inertia, frozen arms and Jacobians, friction) adds it.
//
// Steps are reversed from the last; the body state is checkpointed every k
// steps and the states inside a segment are recomputed by the same forward
// replay (bit-exact with the engine's steps).
#define MX_UNROLL_POINTS _Pragma("unroll")
#include "mx_diffrec.h"

#include <algorithm>
#include <string.h>
#include <vector>

// Workgroup lane, lane count and barrier; the host build runs one world with
// a single lane (every per-lane loop serial, barriers empty), which is the
// same computation.
#if defined(__HIP_DEVICE_COMPILE__)
#define ADJ_TID ((int)threadIdx.x)
#define ADJ_NT ((int)blockDim.x)
#define ADJ_SYNC() __syncthreads()
#else
#define ADJ_TID 0
#define ADJ_NT 1
#define ADJ_SYNC() ((void)0)
#endif

namespace mx
{

using dd::ContactGeo;
using dd::Dual;
using dd::M3;
using dd::PointPrep;
using dd::Q4;
using dd::V3;
typedef Dual<float, double> DF;
typedef dd::Body<DF> DBody;
typedef dd::Body<float> FBody;

constexpr int kAdjThreads = 256;
constexpr int kAdjTasks = 2048;      // results of one chunk of (op, component) tasks
constexpr int kAdjMaxContacts = 3072; // per world (level lists live in LDS)
constexpr int kAdjMaxBodies = 2048;
// a workgroup's static LDS stays within 64 KiB
static_assert(kAdjTasks * 8 + (3 * kAdjMaxContacts + 2 + kAdjMaxBodies + 1) * 4 <= 65536, "adjoint LDS");

// per-op input components
constexpr int kDimIntegrate = 16;   // x q v w (13), inverse local inertia (3)
constexpr int kDimPrepare = 28;     // qa qb, iwA iwB, invMass a b
constexpr int kDimNormals = 44;     // xa qa xb qb, invMass a b, per point wN jA jB
constexpr int kDimFriction = 39;    // xa qa xb qb, iwA iwB, invMass a b, mu, per point lambdaN
constexpr int kDimReconstruct = 14; // x q, previous x q
constexpr int kDimVelocity = 109;   // qa va wa qb vb wb, pv pw a b, iwA iwB, invMass a b, mu, per point 14

// One world's slab (floats), in per-entity records:
//   st     body state x(3) q(4) v(3) w(3)           13 per body
//   rbPre  [s] state at the substep start            13 per body
//   rbPost [s] state after integration               13 per body
//   rbIw   [s] world inverse inertia (columns)        9 per body
//   rbFric [s] x, q after the stiction pass           7 per body
//   rcPrep [s] per point rA rB jA jB wN              52 per contact
//   rcNrm  [s][it] xa qa xb qb, lambdaN before       18 per contact
//   rcLamN [s] lambdaN after the normal iterations    4 per contact
//   rcFric [s] xa qa xb qb before the stiction op    14 per contact
//   rcVel  [s] qa va wa qb vb wb before the op       20 per contact
//   a*     adjoints of st, the previous-substep copies (px pq pv pw), iw,
//          invMass, inverse local inertia, contact friction, lambdaN, and
//          the frozen point data
struct AdjLayout
{
	int32_t MB, MC, S, I;
	int64_t st, im, il, mu, lamN;
	int64_t rbPre, rbPost, rbIw, rbFric;
	int64_t rcPrep, rcNrm, rcLamN, rcFric, rcVel;
	int64_t total;                                      // floats
	int64_t aSt, aPrev, aIw, aIm, aIl, aMu, aLamN, aPrep; // in the double slab
	int64_t dtotal;                                     // doubles
};

static AdjLayout adj_layout(int32_t MB, int32_t MC, int S, int I)
{
	AdjLayout L;
	L.MB = MB;
	L.MC = MC;
	L.S = S;
	L.I = I;
	int64_t o = 0;
	auto take = [&](int64_t n) {
		int64_t r = o;
		o += (n + 31) & ~(int64_t)31;
		return r;
	};
	int64_t B = MB, C = MC;
	L.st = take(B * 13);
	L.im = take(B);
	L.il = take(B * 3);
	L.mu = take(C);
	L.lamN = take(C * 4);
	L.rbPre = take(S * B * 13);
	L.rbPost = take(S * B * 13);
	L.rbIw = take(S * B * 9);
	L.rbFric = take(S * B * 7);
	L.rcPrep = take(S * C * 52);
	L.rcNrm = take((int64_t)S * I * C * 18);
	L.rcLamN = take(S * C * 4);
	L.rcFric = take(S * C * 14);
	L.rcVel = take(S * C * 20);
	L.total = o;
	o = 0;
	L.aSt = take(B * 13);
	L.aPrev = take(B * 13);
	L.aIw = take(B * 9);
	L.aIm = take(B);
	L.aIl = take(B * 3);
	L.aMu = take(C);
	L.aLamN = take(C * 4);
	L.aPrep = take(C * 52);
	L.dtotal = o;
	return L;
}

// ------------------------------------------------------------ helpers

MX_HD V3<float> ld3(const float* p) { return dd::v3<float>(p[0], p[1], p[2]); }
MX_HD Q4<float> ld4(const float* p)
{
	Q4<float> q = { p[0], p[1], p[2], p[3] };
	return q;
}
MX_HD M3<float> ldm(const float* p)
{
	M3<float> m;
	m.cx = ld3(p);
	m.cy = ld3(p + 3);
	m.cz = ld3(p + 6);
	return m;
}
MX_HD void st3(float* p, V3<float> v)
{
	p[0] = v.x;
	p[1] = v.y;
	p[2] = v.z;
}
MX_HD void st4(float* p, Q4<float> q)
{
	p[0] = q.x;
	p[1] = q.y;
	p[2] = q.z;
	p[3] = q.w;
}
MX_HD void stm(float* p, const M3<float>& m)
{
	st3(p, m.cx);
	st3(p + 3, m.cy);
	st3(p + 6, m.cz);
}

// component k of an op's inputs, with tangent 1 when k is the evaluated one
MX_HD DF dv(float v, int k, int d) { return dd::dual<float, double>(v, k == d ? 1.0 : 0.0); }
MX_HD DF dc(float v) { return dd::dual<float, double>(v, 0.0); }
MX_HD V3<DF> dv3(const float* p, int k, int d)
{
	return dd::v3<DF>(dv(p[0], k, d), dv(p[1], k + 1, d), dv(p[2], k + 2, d));
}
MX_HD V3<DF> dc3(const float* p) { return dd::v3<DF>(dc(p[0]), dc(p[1]), dc(p[2])); }
MX_HD Q4<DF> dq4(const float* p, int k, int d)
{
	Q4<DF> q = { dv(p[0], k, d), dv(p[1], k + 1, d), dv(p[2], k + 2, d), dv(p[3], k + 3, d) };
	return q;
}
MX_HD M3<DF> dm3(const float* p, int k, int d)
{
	M3<DF> m;
	m.cx = dv3(p, k, d);
	m.cy = dv3(p + 3, k + 3, d);
	m.cz = dv3(p + 6, k + 6, d);
	return m;
}
// adjoint . tangent
MX_HD double td3(const double* a, const V3<DF>& v) { return a[0] * v.x.d + a[1] * v.y.d + a[2] * v.z.d; }
MX_HD double td4(const double* a, const Q4<DF>& q)
{
	return a[0] * q.x.d + a[1] * q.y.d + a[2] * q.z.d + a[3] * q.w.d;
}
MX_HD double tdm(const double* a, const M3<DF>& m)
{
	return td3(a, m.cx) + td3(a + 3, m.cy) + td3(a + 6, m.cz);
}

// One world's slab and level lists.
struct AW
{
	float *st, *im, *il, *mu, *lamN, *rbPre, *rbPost, *rbIw, *rbFric, *rcPrep, *rcNrm, *rcLamN, *rcFric, *rcVel;
	double *aSt, *aPrev, *aIw, *aIm, *aIl, *aMu, *aLamN, *aPrep;
	const int32_t* order; // contacts by level (LDS)
	const int32_t* start; // level starts (LDS)
	int32_t nL;
	int32_t MB, MC, I;
};

struct SubConst
{
	float h, invH, maxCorrection, gravityLen;
	int s;
};

// ------------------------------------------------------------ forward

MX_HD void fwd_substep(const AW& A, const StepIn& ins, const SubConst& K)
{
	const int tid = ADJ_TID, T = ADJ_NT;
	const int32_t nb = ins.nb, nc = ins.nc, s = K.s;
	float* pre = A.rbPre + (int64_t)s * A.MB * 13;
	float* post = A.rbPost + (int64_t)s * A.MB * 13;
	float* iwr = A.rbIw + (int64_t)s * A.MB * 9;
	float* fr = A.rbFric + (int64_t)s * A.MB * 7;
	float* prep = A.rcPrep + (int64_t)s * A.MC * 52;
	float* lamRec = A.rcLamN + (int64_t)s * A.MC * 4;
	// integrate
	for (int32_t i = tid; i < nb; i += T)
	{
		float* x = A.st + i * 13;
		for (int k = 0; k < 13; ++k)
			pre[i * 13 + k] = x[k];
		V3<float> p = ld3(x), v = ld3(x + 7), w = ld3(x + 10);
		Q4<float> q = ld4(x + 3);
		V3<float> il = ld3(A.il + i * 3);
		M3<float> iw;
		if (dd::is_dynamic(ins, i))
		{
			dd::integrate_op(p, q, v, w, il, iw, K.h, ins.vel[i].w, ins.invIL[i].w, ins.angDamp[i], ins.gravity);
			st3(x, p);
			st4(x + 3, q);
			st3(x + 7, v);
			st3(x + 10, w);
		}
		else
			iw = dd::inv_inertia_world(q, il);
		for (int k = 0; k < 13; ++k)
			post[i * 13 + k] = x[k];
		stm(iwr + i * 9, iw);
	}
	ADJ_SYNC();
	// prepare
	for (int32_t c = tid; c < nc; c += T)
	{
		ContactGeo g = dd::contact_geo(ins, c);
		int2 ab = ins.bodies[c];
		Q4<float> qa = ld4(post + ab.x * 13 + 3), qb = ld4(post + ab.y * 13 + 3);
		M3<float> iwa = ldm(iwr + ab.x * 9), iwb = ldm(iwr + ab.y * 9);
		for (int j = 0; j < g.count; ++j)
		{
			PointPrep<float> r = dd::prepare_point(g, j, qa, qb, iwa, iwb, A.im[ab.x], A.im[ab.y]);
			float* o = prep + c * 52 + j * 13;
			st3(o, r.rA);
			st3(o + 3, r.rB);
			st3(o + 6, r.jA);
			st3(o + 9, r.jB);
			o[12] = r.wN;
			A.lamN[c * 4 + j] = 0.0f;
		}
	}
	ADJ_SYNC();
	// normal projection iterations, level by level
	for (int it = 0; it < A.I; ++it)
	{
		float* rec = A.rcNrm + ((int64_t)s * A.I + it) * A.MC * 18;
		for (int32_t l = 0; l < A.nL; ++l)
		{
			for (int32_t k = A.start[l] + tid; k < A.start[l + 1]; k += T)
			{
				int32_t c = A.order[k];
				if ((uint32_t)c >= (uint32_t)nc)
					continue;
				ContactGeo g = dd::contact_geo(ins, c);
				int2 ab = ins.bodies[c];
				float* xa = A.st + ab.x * 13;
				float* xb = A.st + ab.y * 13;
				V3<float> pa = ld3(xa), pb = ld3(xb);
				Q4<float> qa = ld4(xa + 3), qb = ld4(xb + 3);
				float lN[4], wN[4];
				V3<float> jA[4], jB[4];
				const float* pr = prep + c * 52;
				for (int j = 0; j < 4; ++j)
				{
					lN[j] = A.lamN[c * 4 + j];
					wN[j] = pr[j * 13 + 12];
					jA[j] = ld3(pr + j * 13 + 6);
					jB[j] = ld3(pr + j * 13 + 9);
				}
				float* r = rec + c * 18;
				st3(r, pa);
				st4(r + 3, qa);
				st3(r + 7, pb);
				st4(r + 10, qb);
				for (int j = 0; j < 4; ++j)
					r[14 + j] = lN[j];
				dd::normals_any(ins, c, g, pa, qa, pb, qb, A.im[ab.x], A.im[ab.y], lN, wN, jA, jB, K.maxCorrection);
				for (int j = 0; j < g.count; ++j)
					A.lamN[c * 4 + j] = lN[j];
				if (g.dynA)
				{
					st3(xa, pa);
					st4(xa + 3, qa);
				}
				if (g.dynB)
				{
					st3(xb, pb);
					st4(xb + 3, qb);
				}
			}
			ADJ_SYNC();
		}
	}
	for (int32_t c = tid; c < nc; c += T)
		for (int j = 0; j < 4; ++j)
			lamRec[c * 4 + j] = A.lamN[c * 4 + j];
	// stiction (box-box pairs), level by level
	{
		float* rec = A.rcFric + (int64_t)s * A.MC * 14;
		for (int32_t l = 0; l < A.nL; ++l)
		{
			for (int32_t k = A.start[l] + tid; k < A.start[l + 1]; k += T)
			{
				int32_t c = A.order[k];
				if ((uint32_t)c >= (uint32_t)nc)
					continue;
				ContactGeo g = dd::contact_geo(ins, c);
				if (g.tA != M3D_SHAPE_BOX || g.tB != M3D_SHAPE_BOX || g.count == 0)
					continue;
				int2 ab = ins.bodies[c];
				float* xa = A.st + ab.x * 13;
				float* xb = A.st + ab.y * 13;
				FBody a, b;
				a.p = ld3(xa);
				a.q = ld4(xa + 3);
				a.v = ld3(xa + 7);
				a.w = ld3(xa + 10);
				a.invMass = A.im[ab.x];
				a.iw = ldm(iwr + ab.x * 9);
				b.p = ld3(xb);
				b.q = ld4(xb + 3);
				b.v = ld3(xb + 7);
				b.w = ld3(xb + 10);
				b.invMass = A.im[ab.y];
				b.iw = ldm(iwr + ab.y * 9);
				float* r = rec + c * 14;
				st3(r, a.p);
				st4(r + 3, a.q);
				st3(r + 7, b.p);
				st4(r + 10, b.q);
				float lN[4], lT[4];
				float4 fA[4], fB[4];
				for (int j = 0; j < 4; ++j)
				{
					lN[j] = A.lamN[c * 4 + j];
					lT[j] = 0.0f;
					fA[j] = ins.pFA[c * 4 + j];
					fB[j] = ins.pFB[c * 4 + j];
				}
				const float* pr = prep + c * 52;
				dd::friction_op(g, a, b, A.mu[c], lN, lT, ld3(pr), ld3(pr + 3), fA, fB);
				if (g.dynA)
				{
					st3(xa, a.p);
					st4(xa + 3, a.q);
				}
				if (g.dynB)
				{
					st3(xb, b.p);
					st4(xb + 3, b.q);
				}
			}
			ADJ_SYNC();
		}
	}
	// velocities from positions
	for (int32_t i = tid; i < nb; i += T)
	{
		float* x = A.st + i * 13;
		for (int k = 0; k < 7; ++k)
			fr[i * 7 + k] = x[k];
		if (!dd::is_dynamic(ins, i))
			continue;
		V3<float> v, w;
		dd::reconstruct_op(ld3(x), ld4(x + 3), ld3(pre + i * 13), ld4(pre + i * 13 + 3), K.invH, v, w);
		st3(x + 7, v);
		st3(x + 10, w);
	}
	ADJ_SYNC();
	// velocity pass, level by level
	{
		float* rec = A.rcVel + (int64_t)s * A.MC * 20;
		for (int32_t l = 0; l < A.nL; ++l)
		{
			for (int32_t k = A.start[l] + tid; k < A.start[l + 1]; k += T)
			{
				int32_t c = A.order[k];
				if ((uint32_t)c >= (uint32_t)nc)
					continue;
				ContactGeo g = dd::contact_geo(ins, c);
				int2 ab = ins.bodies[c];
				float* xa = A.st + ab.x * 13;
				float* xb = A.st + ab.y * 13;
				FBody a, b;
				a.p = ld3(xa);
				a.q = ld4(xa + 3);
				a.v = ld3(xa + 7);
				a.w = ld3(xa + 10);
				a.invMass = A.im[ab.x];
				a.iw = ldm(iwr + ab.x * 9);
				b.p = ld3(xb);
				b.q = ld4(xb + 3);
				b.v = ld3(xb + 7);
				b.w = ld3(xb + 10);
				b.invMass = A.im[ab.y];
				b.iw = ldm(iwr + ab.y * 9);
				float* r = rec + c * 20;
				st4(r, a.q);
				st3(r + 4, a.v);
				st3(r + 7, a.w);
				st4(r + 10, b.q);
				st3(r + 14, b.v);
				st3(r + 17, b.w);
				float lN[4];
				PointPrep<float> pp[4];
				float4 slip[4];
				const float* pr = prep + c * 52;
				for (int j = 0; j < 4; ++j)
				{
					lN[j] = A.lamN[c * 4 + j];
					pp[j].rA = ld3(pr + j * 13);
					pp[j].rB = ld3(pr + j * 13 + 3);
					pp[j].jA = ld3(pr + j * 13 + 6);
					pp[j].jB = ld3(pr + j * 13 + 9);
					pp[j].wN = pr[j * 13 + 12];
					slip[j] = ins.pSlip[c * 4 + j];
				}
				dd::velocity_op(g, a, b, ld3(pre + ab.x * 13 + 7), ld3(pre + ab.x * 13 + 10), ld3(pre + ab.y * 13 + 7),
								ld3(pre + ab.y * 13 + 10), A.mu[c], lN, pp, slip, K.h, K.gravityLen);
				if (g.dynA)
				{
					st3(xa + 7, a.v);
					st3(xa + 10, a.w);
				}
				if (g.dynB)
				{
					st3(xb + 7, b.v);
					st3(xb + 10, b.w);
				}
			}
			ADJ_SYNC();
		}
	}
}

// ------------------------------------------------------------ op products

// Velocity pass op of contact c, component d (kDimVelocity layout:
// 0-9 qa va wa, 10-19 qb vb wb, 20-25 pvA pwA, 26-31 pvB pwB, 32-40 iwA,
// 41-49 iwB, 50-51 invMass a b, 52 mu, 53 + 14 j + m point j: m 0 lambdaN,
// 1 wN, 2-4 rA, 5-7 rB, 8-10 jA, 11-13 jB).
MX_HD bool vel_active(const ContactGeo& g, int d)
{
	bool curved = (f2u(g.mat.w) & CF_CURVED) != 0;
	if (d < 53)
	{
		bool onA = d < 10 || (d >= 20 && d < 26) || (d >= 32 && d < 41) || d == 50;
		bool onB = (d >= 10 && d < 20) || (d >= 26 && d < 32) || (d >= 41 && d < 50) || d == 51;
		if ((onA && !g.dynA) || (onB && !g.dynB))
			return false;
		if ((d < 4 || (d >= 10 && d < 14)) && !curved)
			return false;
		return true;
	}
	int j = (d - 53) / 14, m = (d - 53) % 14;
	if (j >= g.count)
		return false;
	if (m >= 2 && m < 8 && curved)
		return false;
	return true;
}

MX_HD double vel_eval(const AW& A, const StepIn& ins, const SubConst& K, int32_t c, int d)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (!vel_active(g, d))
		return 0.0;
	const int32_t s = K.s;
	int2 ab = ins.bodies[c];
	const float* rv = A.rcVel + ((int64_t)s * A.MC + c) * 20;
	const float* pre = A.rbPre + (int64_t)s * A.MB * 13;
	const float* iwr = A.rbIw + (int64_t)s * A.MB * 9;
	const float* pr = A.rcPrep + ((int64_t)s * A.MC + c) * 52;
	const float* lr = A.rcLamN + ((int64_t)s * A.MC + c) * 4;
	DBody a, b;
	a.p = dd::v3<DF>(dc(0.0f), dc(0.0f), dc(0.0f));
	b.p = a.p;
	a.q = dq4(rv, 0, d);
	a.v = dv3(rv + 4, 4, d);
	a.w = dv3(rv + 7, 7, d);
	b.q = dq4(rv + 10, 10, d);
	b.v = dv3(rv + 14, 14, d);
	b.w = dv3(rv + 17, 17, d);
	V3<DF> pvA = dv3(pre + ab.x * 13 + 7, 20, d), pwA = dv3(pre + ab.x * 13 + 10, 23, d);
	V3<DF> pvB = dv3(pre + ab.y * 13 + 7, 26, d), pwB = dv3(pre + ab.y * 13 + 10, 29, d);
	a.iw = dm3(iwr + ab.x * 9, 32, d);
	b.iw = dm3(iwr + ab.y * 9, 41, d);
	a.invMass = dv(A.im[ab.x], 50, d);
	b.invMass = dv(A.im[ab.y], 51, d);
	DF mu = dv(A.mu[c], 52, d);
	DF lN[4];
	PointPrep<DF> pp[4];
	float4 slip[4];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		int k = 53 + 14 * j;
		lN[j] = dv(lr[j], k, d);
		pp[j].wN = dv(pr[j * 13 + 12], k + 1, d);
		pp[j].rA = dv3(pr + j * 13, k + 2, d);
		pp[j].rB = dv3(pr + j * 13 + 3, k + 5, d);
		pp[j].jA = dv3(pr + j * 13 + 6, k + 8, d);
		pp[j].jB = dv3(pr + j * 13 + 9, k + 11, d);
		slip[j] = ins.pSlip[c * 4 + j];
	}
	dd::velocity_op(g, a, b, pvA, pwA, pvB, pwB, mu, lN, pp, slip, K.h, K.gravityLen);
	double r = 0.0;
	if (g.dynA)
		r += td3(A.aSt + ab.x * 13 + 7, a.v) + td3(A.aSt + ab.x * 13 + 10, a.w);
	if (g.dynB)
		r += td3(A.aSt + ab.y * 13 + 7, b.v) + td3(A.aSt + ab.y * 13 + 10, b.w);
	return r;
}

MX_HD void vel_apply(const AW& A, const StepIn& ins, int32_t c, int d, double r)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (!vel_active(g, d))
		return;
	int2 ab = ins.bodies[c];
	if (d < 10)
	{
		if (d < 4)
			A.aSt[ab.x * 13 + 3 + d] += r;
		else
			A.aSt[ab.x * 13 + 7 + (d - 4)] = r;
	}
	else if (d < 20)
	{
		if (d < 14)
			A.aSt[ab.y * 13 + 3 + (d - 10)] += r;
		else
			A.aSt[ab.y * 13 + 7 + (d - 14)] = r;
	}
	else if (d < 26)
		A.aPrev[ab.x * 13 + 7 + (d - 20)] += r;
	else if (d < 32)
		A.aPrev[ab.y * 13 + 7 + (d - 26)] += r;
	else if (d < 41)
		A.aIw[ab.x * 9 + (d - 32)] += r;
	else if (d < 50)
		A.aIw[ab.y * 9 + (d - 41)] += r;
	else if (d == 50)
		A.aIm[ab.x] += r;
	else if (d == 51)
		A.aIm[ab.y] += r;
	else if (d == 52)
		A.aMu[c] += r;
	else
	{
		int j = (d - 53) / 14, m = (d - 53) % 14;
		if (m == 0)
			A.aLamN[c * 4 + j] += r;
		else
			A.aPrep[c * 52 + j * 13 + (m == 1 ? 12 : m - 2)] += r;
	}
}

// Velocity reconstruction of body i (kDimReconstruct: 0-6 x q, 7-13
// previous x q).
MX_HD double rec_eval(const AW& A, const StepIn& ins, const SubConst& K, int32_t i, int d)
{
	if (!dd::is_dynamic(ins, i))
		return 0.0;
	const float* fr = A.rbFric + ((int64_t)K.s * A.MB + i) * 7;
	const float* pre = A.rbPre + ((int64_t)K.s * A.MB + i) * 13;
	V3<DF> v, w;
	dd::reconstruct_op(dv3(fr, 0, d), dq4(fr + 3, 3, d), dv3(pre, 7, d), dq4(pre + 3, 10, d), K.invH, v, w);
	return td3(A.aSt + i * 13 + 7, v) + td3(A.aSt + i * 13 + 10, w);
}

MX_HD void rec_apply(const AW& A, const StepIn& ins, int32_t i, int d, double r)
{
	if (!dd::is_dynamic(ins, i))
		return;
	if (d < 7)
		A.aSt[i * 13 + d] += r;
	else
		A.aPrev[i * 13 + (d - 7)] += r;
	if (d == 0)
		for (int k = 7; k < 13; ++k)
			A.aSt[i * 13 + k] = 0.0; // overwritten by the reconstruction
}

// Stiction op of contact c (kDimFriction: 0-6 xa qa, 7-13 xb qb, 14-22 iwA,
// 23-31 iwB, 32-33 invMass a b, 34 mu, 35 + j lambdaN of point j).
MX_HD bool fric_active(const ContactGeo& g, int d)
{
	bool onA = d < 7 || (d >= 14 && d < 23) || d == 32;
	bool onB = (d >= 7 && d < 14) || (d >= 23 && d < 32) || d == 33;
	if ((onA && !g.dynA) || (onB && !g.dynB))
		return false;
	if (d >= 35 && d - 35 >= g.count)
		return false;
	return true;
}

MX_HD double fric_eval(const AW& A, const StepIn& ins, const SubConst& K, int32_t c, int d)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (!fric_active(g, d))
		return 0.0;
	const int32_t s = K.s;
	int2 ab = ins.bodies[c];
	const float* rf = A.rcFric + ((int64_t)s * A.MC + c) * 14;
	const float* post = A.rbPost + (int64_t)s * A.MB * 13;
	const float* iwr = A.rbIw + (int64_t)s * A.MB * 9;
	const float* pr = A.rcPrep + ((int64_t)s * A.MC + c) * 52;
	const float* lr = A.rcLamN + ((int64_t)s * A.MC + c) * 4;
	DBody a, b;
	a.p = dv3(rf, 0, d);
	a.q = dq4(rf + 3, 3, d);
	b.p = dv3(rf + 7, 7, d);
	b.q = dq4(rf + 10, 10, d);
	a.v = dc3(post + ab.x * 13 + 7);
	a.w = dc3(post + ab.x * 13 + 10);
	b.v = dc3(post + ab.y * 13 + 7);
	b.w = dc3(post + ab.y * 13 + 10);
	a.iw = dm3(iwr + ab.x * 9, 14, d);
	b.iw = dm3(iwr + ab.y * 9, 23, d);
	a.invMass = dv(A.im[ab.x], 32, d);
	b.invMass = dv(A.im[ab.y], 33, d);
	DF mu = dv(A.mu[c], 34, d);
	DF lN[4], lT[4];
	float4 fA[4], fB[4];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		lN[j] = dv(lr[j], 35 + j, d);
		lT[j] = dc(0.0f);
		fA[j] = ins.pFA[c * 4 + j];
		fB[j] = ins.pFB[c * 4 + j];
	}
	dd::friction_op(g, a, b, mu, lN, lT, dc3(pr), dc3(pr + 3), fA, fB);
	double r = 0.0;
	if (g.dynA)
		r += td3(A.aSt + ab.x * 13, a.p) + td4(A.aSt + ab.x * 13 + 3, a.q);
	if (g.dynB)
		r += td3(A.aSt + ab.y * 13, b.p) + td4(A.aSt + ab.y * 13 + 3, b.q);
	return r;
}

MX_HD void fric_apply(const AW& A, const StepIn& ins, int32_t c, int d, double r)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (!fric_active(g, d))
		return;
	int2 ab = ins.bodies[c];
	if (d < 7)
		A.aSt[ab.x * 13 + d] = r;
	else if (d < 14)
		A.aSt[ab.y * 13 + (d - 7)] = r;
	else if (d < 23)
		A.aIw[ab.x * 9 + (d - 14)] += r;
	else if (d < 32)
		A.aIw[ab.y * 9 + (d - 23)] += r;
	else if (d == 32)
		A.aIm[ab.x] += r;
	else if (d == 33)
		A.aIm[ab.y] += r;
	else if (d == 34)
		A.aMu[c] += r;
	else
		A.aLamN[c * 4 + (d - 35)] += r;
}

// Normal projection op of contact c in iteration it (kDimNormals: 0-6 xa
// qa, 7-13 xb qb, 14-15 invMass a b, 16 + 7 j + m point j: m 0 wN, 1-3 jA,
// 4-6 jB). lambdaN is carried through unchanged (the op adds to it), so its
// adjoint passes through and only enters through the increments.
MX_HD bool nrm_active(const ContactGeo& g, int d)
{
	bool onA = d < 7 || d == 14;
	bool onB = (d >= 7 && d < 14) || d == 15;
	if ((onA && !g.dynA) || (onB && !g.dynB))
		return false;
	if (d >= 16 && (d - 16) / 7 >= g.count)
		return false;
	return true;
}

MX_HD double nrm_eval(const AW& A, const StepIn& ins, const SubConst& K, int it, int32_t c, int d)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (!nrm_active(g, d))
		return 0.0;
	const int32_t s = K.s;
	int2 ab = ins.bodies[c];
	const float* rn = A.rcNrm + (((int64_t)s * A.I + it) * A.MC + c) * 18;
	const float* pr = A.rcPrep + ((int64_t)s * A.MC + c) * 52;
	V3<DF> pa = dv3(rn, 0, d), pb = dv3(rn + 7, 7, d);
	Q4<DF> qa = dq4(rn + 3, 3, d), qb = dq4(rn + 10, 10, d);
	DF imA = dv(A.im[ab.x], 14, d), imB = dv(A.im[ab.y], 15, d);
	DF lN[4], wN[4];
	V3<DF> jA[4], jB[4];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		int k = 16 + 7 * j;
		lN[j] = dc(rn[14 + j]);
		wN[j] = dv(pr[j * 13 + 12], k, d);
		jA[j] = dv3(pr + j * 13 + 6, k + 1, d);
		jB[j] = dv3(pr + j * 13 + 9, k + 4, d);
	}
	dd::normals_any(ins, c, g, pa, qa, pb, qb, imA, imB, lN, wN, jA, jB, K.maxCorrection);
	double r = 0.0;
	if (g.dynA)
		r += td3(A.aSt + ab.x * 13, pa) + td4(A.aSt + ab.x * 13 + 3, qa);
	if (g.dynB)
		r += td3(A.aSt + ab.y * 13, pb) + td4(A.aSt + ab.y * 13 + 3, qb);
#pragma unroll
	for (int j = 0; j < 4; ++j)
		if (j < g.count)
			r += A.aLamN[c * 4 + j] * lN[j].d;
	return r;
}

MX_HD void nrm_apply(const AW& A, const StepIn& ins, int32_t c, int d, double r)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (!nrm_active(g, d))
		return;
	int2 ab = ins.bodies[c];
	if (d < 7)
		A.aSt[ab.x * 13 + d] = r;
	else if (d < 14)
		A.aSt[ab.y * 13 + (d - 7)] = r;
	else if (d == 14)
		A.aIm[ab.x] += r;
	else if (d == 15)
		A.aIm[ab.y] += r;
	else
	{
		int j = (d - 16) / 7, m = (d - 16) % 7;
		A.aPrep[c * 52 + j * 13 + (m == 0 ? 12 : 5 + m)] += r; // wN at 12, jA at 6-8, jB at 9-11
	}
}

// Prepare op of contact c (kDimPrepare: 0-3 qa, 4-7 qb, 8-16 iwA, 17-25 iwB,
// 26-27 invMass a b).
MX_HD bool prep_active(const ContactGeo& g, int d)
{
	bool onA = d < 4 || (d >= 8 && d < 17) || d == 26;
	return onA ? g.dynA : g.dynB;
}

MX_HD double prep_eval(const AW& A, const StepIn& ins, const SubConst& K, int32_t c, int d)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (g.count == 0 || !prep_active(g, d))
		return 0.0;
	const int32_t s = K.s;
	int2 ab = ins.bodies[c];
	const float* post = A.rbPost + (int64_t)s * A.MB * 13;
	const float* iwr = A.rbIw + (int64_t)s * A.MB * 9;
	Q4<DF> qa = dq4(post + ab.x * 13 + 3, 0, d), qb = dq4(post + ab.y * 13 + 3, 4, d);
	M3<DF> iwa = dm3(iwr + ab.x * 9, 8, d), iwb = dm3(iwr + ab.y * 9, 17, d);
	DF imA = dv(A.im[ab.x], 26, d), imB = dv(A.im[ab.y], 27, d);
	double r = 0.0;
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		if (j >= g.count)
			break;
		PointPrep<DF> p = dd::prepare_point(g, j, qa, qb, iwa, iwb, imA, imB);
		const double* ap = A.aPrep + c * 52 + j * 13;
		r += td3(ap, p.rA) + td3(ap + 3, p.rB) + td3(ap + 6, p.jA) + td3(ap + 9, p.jB) + ap[12] * p.wN.d;
	}
	return r;
}

MX_HD void prep_apply(const AW& A, const StepIn& ins, int32_t c, int d, double r)
{
	ContactGeo g = dd::contact_geo(ins, c);
	if (g.count == 0 || !prep_active(g, d))
		return;
	int2 ab = ins.bodies[c];
	if (d < 4)
		A.aSt[ab.x * 13 + 3 + d] += r;
	else if (d < 8)
		A.aSt[ab.y * 13 + 3 + (d - 4)] += r;
	else if (d < 17)
		A.aIw[ab.x * 9 + (d - 8)] += r;
	else if (d < 26)
		A.aIw[ab.y * 9 + (d - 17)] += r;
	else if (d == 26)
		A.aIm[ab.x] += r;
	else
		A.aIm[ab.y] += r;
}

// Integration of body i (kDimIntegrate: 0-12 x q v w at the substep start,
// 13-15 inverse local inertia). Outputs: the state, the previous-substep
// copies (the inputs themselves) and the world inverse inertia.
MX_HD double int_eval(const AW& A, const StepIn& ins, const SubConst& K, int32_t i, int d)
{
	if (!dd::is_dynamic(ins, i))
		return 0.0;
	const float* pre = A.rbPre + ((int64_t)K.s * A.MB + i) * 13;
	V3<DF> x = dv3(pre, 0, d), v = dv3(pre + 7, 7, d), w = dv3(pre + 10, 10, d);
	Q4<DF> q = dq4(pre + 3, 3, d);
	V3<DF> il = dv3(A.il + i * 3, 13, d);
	M3<DF> iw;
	dd::integrate_op(x, q, v, w, il, iw, K.h, ins.vel[i].w, ins.invIL[i].w, ins.angDamp[i], ins.gravity);
	const double* as = A.aSt + i * 13;
	double r = td3(as, x) + td4(as + 3, q) + td3(as + 7, v) + td3(as + 10, w) + tdm(A.aIw + i * 9, iw);
	if (d < 13)
		r += A.aPrev[i * 13 + d];
	return r;
}

MX_HD void int_apply(const AW& A, const StepIn& ins, int32_t i, int d, double r)
{
	if (!dd::is_dynamic(ins, i))
		return;
	if (d < 13)
		A.aSt[i * 13 + d] = r;
	else
		A.aIl[i * 3 + (d - 13)] += r;
}

// Runs the (item, component) products of items list[k0 .. k1) (list null:
// items k0 .. k1), reading everything before writing anything.
template <class Eval, class Apply>
MX_HD void run_products(double* res, const int32_t* list, int32_t k0, int32_t k1, int32_t bound, int D, Eval ev,
						Apply ap)
{
	const int32_t per = kAdjTasks / D;
	for (int32_t base = k0; base < k1; base += per)
	{
		int32_t e = base + per < k1 ? base + per : k1;
		int32_t n = (e - base) * D;
		for (int32_t t = ADJ_TID; t < n; t += ADJ_NT)
		{
			int32_t k = base + t / D;
			int32_t item = list != nullptr ? list[k] : k;
			res[t] = (uint32_t)item < (uint32_t)bound ? ev(item, t % D) : 0.0;
		}
		ADJ_SYNC();
		for (int32_t t = ADJ_TID; t < n; t += ADJ_NT)
		{
			int32_t k = base + t / D;
			int32_t item = list != nullptr ? list[k] : k;
			if ((uint32_t)item < (uint32_t)bound)
				ap(item, t % D, res[t]);
		}
		ADJ_SYNC();
	}
}

MX_HD void rev_substep(const AW& A, const StepIn& ins, const SubConst& K, double* res)
{
	const int tid = ADJ_TID, T = ADJ_NT;
	const int32_t nb = ins.nb, nc = ins.nc;
	for (int32_t k = tid; k < nb * 13; k += T)
		A.aPrev[k] = 0.0;
	for (int32_t k = tid; k < nb * 9; k += T)
		A.aIw[k] = 0.0;
	for (int32_t k = tid; k < nc * 4; k += T)
		A.aLamN[k] = 0.0;
	for (int32_t k = tid; k < nc * 52; k += T)
		A.aPrep[k] = 0.0;
	ADJ_SYNC();
	// velocity pass
	for (int32_t l = A.nL - 1; l >= 0; --l)
		run_products(
			res, A.order, A.start[l], A.start[l + 1], nc, kDimVelocity,
			[&](int32_t c, int d) { return vel_eval(A, ins, K, c, d); },
			[&](int32_t c, int d, double r) { vel_apply(A, ins, c, d, r); });
	// velocity reconstruction
	run_products(
		res, nullptr, 0, nb, nb, kDimReconstruct, [&](int32_t i, int d) { return rec_eval(A, ins, K, i, d); },
		[&](int32_t i, int d, double r) { rec_apply(A, ins, i, d, r); });
	// stiction (box-box pairs; the others' products are zero and not applied)
	for (int32_t l = A.nL - 1; l >= 0; --l)
		run_products(
			res, A.order, A.start[l], A.start[l + 1], nc, kDimFriction,
			[&](int32_t c, int d) {
				ContactGeo g = dd::contact_geo(ins, c);
				if (g.tA != M3D_SHAPE_BOX || g.tB != M3D_SHAPE_BOX || g.count == 0)
					return 0.0;
				return fric_eval(A, ins, K, c, d);
			},
			[&](int32_t c, int d, double r) {
				ContactGeo g = dd::contact_geo(ins, c);
				if (g.tA != M3D_SHAPE_BOX || g.tB != M3D_SHAPE_BOX || g.count == 0)
					return;
				fric_apply(A, ins, c, d, r);
			});
	// normal projection iterations
	for (int it = A.I - 1; it >= 0; --it)
		for (int32_t l = A.nL - 1; l >= 0; --l)
			run_products(
				res, A.order, A.start[l], A.start[l + 1], nc, kDimNormals,
				[&](int32_t c, int d) { return nrm_eval(A, ins, K, it, c, d); },
				[&](int32_t c, int d, double r) { nrm_apply(A, ins, c, d, r); });
	// prepare (levels keep contacts that share a body apart)
	for (int32_t l = A.nL - 1; l >= 0; --l)
		run_products(
			res, A.order, A.start[l], A.start[l + 1], nc, kDimPrepare,
			[&](int32_t c, int d) { return prep_eval(A, ins, K, c, d); },
			[&](int32_t c, int d, double r) { prep_apply(A, ins, c, d, r); });
	// integration
	run_products(
		res, nullptr, 0, nb, nb, kDimIntegrate, [&](int32_t i, int d) { return int_eval(A, ins, K, i, d); },
		[&](int32_t i, int d, double r) { int_apply(A, ins, i, d, r); });
}

// ------------------------------------------------------------ kernel

// One recorded step of every world. stateIn: 13 floats per body (x q v w;
// null: the recorded state); stateOut (nullable): the state after the solve.
// reverse != 0: seedIn is the adjoint of the output state, seedOut receives
// the adjoint of the input state, massGrad and fricGrad accumulate the
// step's share of the parameter gradients.
// Workgroup-local lists of one world (LDS on the device).
struct AdjLocal
{
	double* res;     // kAdjTasks
	int32_t* order;  // kAdjMaxContacts
	int32_t* level;  // kAdjMaxContacts
	int32_t* start;  // kAdjMaxContacts + 2
	int32_t* last;   // kAdjMaxBodies
	int32_t* nL;     // 1
};

MX_HD void adj_world(const RecView& V, const AdjLayout& L, float* slab, double* dslab, const float* stateIn,
					 float* stateOut, int reverse, const double* seedIn, double* seedOut, double* massGrad,
					 double* fricGrad, int32_t wi, const AdjLocal& Lc)
{
	double* res = Lc.res;
	int32_t* sOrder = Lc.order;
	int32_t* sLevel = Lc.level;
	int32_t* sStart = Lc.start;
	int32_t* sLast = Lc.last;
	int32_t& sNL = *Lc.nL;
	const int tid = ADJ_TID, T = ADJ_NT;
	const StepIn in = world_in(V, wi);
	const int32_t b0 = V.bodyStart[wi];
	const int32_t nb = in.nb, nc = in.nc;
	float* W = slab + (int64_t)wi * L.total;
	double* WD = dslab + (int64_t)wi * L.dtotal;
	AW A;
	A.st = W + L.st;
	A.im = W + L.im;
	A.il = W + L.il;
	A.mu = W + L.mu;
	A.lamN = W + L.lamN;
	A.rbPre = W + L.rbPre;
	A.rbPost = W + L.rbPost;
	A.rbIw = W + L.rbIw;
	A.rbFric = W + L.rbFric;
	A.rcPrep = W + L.rcPrep;
	A.rcNrm = W + L.rcNrm;
	A.rcLamN = W + L.rcLamN;
	A.rcFric = W + L.rcFric;
	A.rcVel = W + L.rcVel;
	A.aSt = WD + L.aSt;
	A.aPrev = WD + L.aPrev;
	A.aIw = WD + L.aIw;
	A.aIm = WD + L.aIm;
	A.aIl = WD + L.aIl;
	A.aMu = WD + L.aMu;
	A.aLamN = WD + L.aLamN;
	A.aPrep = WD + L.aPrep;
	A.order = sOrder;
	A.start = sStart;
	A.MB = L.MB;
	A.MC = L.MC;
	A.I = L.I;

	// state, parameters and contact friction (src/world.cpp:1554: sqrt of
	// the product of the bodies' friction)
	for (int32_t i = tid; i < nb; i += T)
	{
		float* x = A.st + i * 13;
		if (stateIn != nullptr)
		{
			const float* si = stateIn + (int64_t)(b0 + i) * 13;
			for (int k = 0; k < 13; ++k)
				x[k] = si[k];
		}
		else
		{
			float4 P = in.pos[i], Q = in.rot[i], Vv = in.vel[i], Ww = in.angVel[i];
			x[0] = P.x;
			x[1] = P.y;
			x[2] = P.z;
			x[3] = Q.x;
			x[4] = Q.y;
			x[5] = Q.z;
			x[6] = Q.w;
			x[7] = Vv.x;
			x[8] = Vv.y;
			x[9] = Vv.z;
			x[10] = Ww.x;
			x[11] = Ww.y;
			x[12] = Ww.z;
		}
		A.im[i] = in.pos[i].w;
		float4 IL = in.invIL[i];
		A.il[i * 3] = IL.x;
		A.il[i * 3 + 1] = IL.y;
		A.il[i * 3 + 2] = IL.z;
		sLast[i] = -1;
	}
	// packed pairs for the level pass: a | b << 12 | dynamic bits | the
	// contact's pass-order key << 26
	for (int32_t c = tid; c < nc; c += T)
	{
		int2 ab = in.bodies[c];
		A.mu[c] = dd::sqrt_(in.bmat[ab.x].y * in.bmat[ab.y].y);
		sOrder[c] = ab.x | (ab.y << 12) | (dd::is_dynamic(in, ab.x) ? 1 << 24 : 0) |
					(dd::is_dynamic(in, ab.y) ? 1 << 25 : 0) | (dd::sched_key(in, c) << 26);
	}
	ADJ_SYNC();
	// levels of the step's contact order: key by key (a colored island's
	// colors, then its overflow contacts; every other contact has key 0),
	// ascending contacts within a key
	if (tid == 0)
	{
		int32_t maxL = -1, keys = 1;
		for (int32_t c = 0; c < nc; ++c)
		{
			int32_t kc = ((sOrder[c] >> 26) & 0xF) + 1;
			keys = keys > kc ? keys : kc;
		}
		for (int32_t key = 0; key < keys; ++key)
			for (int32_t c = 0; c < nc; ++c)
			{
				int32_t pk = sOrder[c];
				if (((pk >> 26) & 0xF) != key)
					continue;
				int32_t a = pk & 0xFFF, b = (pk >> 12) & 0xFFF;
				bool da = (pk >> 24) & 1, db = (pk >> 25) & 1;
				int32_t l = 0;
				if (da)
					l = l > sLast[a] + 1 ? l : sLast[a] + 1;
				if (db)
					l = l > sLast[b] + 1 ? l : sLast[b] + 1;
				sLevel[c] = l;
				if (da)
					sLast[a] = l;
				if (db)
					sLast[b] = l;
				maxL = maxL > l ? maxL : l;
			}
		int32_t nL = maxL + 1;
		for (int32_t l = 0; l <= nL; ++l)
			sStart[l] = 0;
		for (int32_t c = 0; c < nc; ++c)
			sStart[sLevel[c] + 1]++;
		for (int32_t l = 0; l < nL; ++l)
			sStart[l + 1] += sStart[l];
		// stable placement, ascending contacts within a level: the starts
		// serve as cursors (each ends at the next level's start) and are
		// shifted back afterwards
		for (int32_t c = 0; c < nc; ++c)
			sOrder[sStart[sLevel[c]]++] = c;
		for (int32_t l = nL; l > 0; --l)
			sStart[l] = sStart[l - 1];
		sStart[0] = 0;
		sNL = nL;
	}
	ADJ_SYNC();
	A.nL = sNL;

	const float h = in.dt / (float)in.substeps;
	SubConst K;
	K.h = h;
	K.invH = 1.0f / h;
	K.maxCorrection = kMaxDepenetrationSpeed * h;
	K.gravityLen = m3d_length(in.gravity);
	for (int s = 0; s < in.substeps; ++s)
	{
		K.s = s;
		fwd_substep(A, dd::substep_in(in, s), K);
		ADJ_SYNC();
	}
	if (stateOut != nullptr)
		for (int32_t k = tid; k < nb * 13; k += T)
			stateOut[(int64_t)b0 * 13 + k] = A.st[k];
	if (!reverse)
		return;

	for (int32_t k = tid; k < nb * 13; k += T)
		A.aSt[k] = seedIn[(int64_t)b0 * 13 + k];
	for (int32_t i = tid; i < nb; i += T)
	{
		A.aIm[i] = 0.0;
		A.aIl[i * 3] = A.aIl[i * 3 + 1] = A.aIl[i * 3 + 2] = 0.0;
	}
	for (int32_t c = tid; c < nc; c += T)
		A.aMu[c] = 0.0;
	ADJ_SYNC();
	for (int s = in.substeps - 1; s >= 0; --s)
	{
		K.s = s;
		rev_substep(A, dd::substep_in(in, s), K, res);
		ADJ_SYNC();
	}
	for (int32_t k = tid; k < nb * 13; k += T)
		seedOut[(int64_t)b0 * 13 + k] = A.aSt[k];
	// mass: the recorded inverse mass and inverse inertia scale as 1 / m
	for (int32_t i = tid; i < nb; i += T)
	{
		float4 P = in.pos[i], IL = in.invIL[i];
		double m = in.angVel[i].w;
		if (dd::is_dynamic(in, i) && P.w != 0.0f && m > 0.0)
			massGrad[b0 + i] += A.aIm[i] * (-(double)P.w / m) + A.aIl[i * 3] * (-(double)IL.x / m) +
								A.aIl[i * 3 + 1] * (-(double)IL.y / m) +
								A.aIl[i * 3 + 2] * (-(double)IL.z / m);
	}
	ADJ_SYNC();
	// friction: mu_c = sqrt(mu_a mu_b), summed into the bodies in contact order
	if (tid == 0)
		for (int32_t c = 0; c < nc; ++c)
		{
			int2 ab = in.bodies[c];
			float ma = in.bmat[ab.x].y, mb = in.bmat[ab.y].y;
			float r = dd::sqrt_(ma * mb);
			double dp = A.aMu[c] * (r > 0.0f ? 0.5 / (double)r : 0.0);
			fricGrad[b0 + ab.x] += dp * (double)mb;
			fricGrad[b0 + ab.y] += dp * (double)ma;
		}
}

// One workgroup per world and step. Reverse != 0: the reverse sweep of the
// step (forward replay with records, then the VJPs); Reverse == 0: the
// forward recomputation of a checkpoint segment's states. Two kernels so that
// a kernel trace tells the two apart.
template <int Reverse>
__global__ __launch_bounds__(kAdjThreads) void k_adj_step(RecView V, AdjLayout L, float* slab, double* dslab,
														  const float* stateIn, float* stateOut,
														  const double* seedIn, double* seedOut, double* massGrad,
														  double* fricGrad)
{
	__shared__ double res[kAdjTasks];
	__shared__ int32_t sOrder[kAdjMaxContacts];
	__shared__ int32_t sLevel[kAdjMaxContacts];
	__shared__ int32_t sStart[kAdjMaxContacts + 2];
	__shared__ int32_t sLast[kAdjMaxBodies];
	__shared__ int32_t sNL;
	const int32_t wi = (int32_t)blockIdx.x;
	if (wi >= V.worlds)
		return;
	AdjLocal Lc = { res, sOrder, sLevel, sStart, sLast, &sNL };
	adj_world(V, L, slab, dslab, stateIn, stateOut, Reverse, seedIn, seedOut, massGrad, fricGrad, wi, Lc);
}

} // namespace mx

// ------------------------------------------------------------ C ABI

using namespace mx;

// Compulsory bytes of one launch: the recorded step (per body 7 float4
// columns, damping and flags; per contact its bodies and, per substep, its
// normal, material and four points of five float4 columns), the state in and
// out, and in reverse the seed in and out and the two gradients.
static double adj_bytes(const DiffStepRec* R, int reverse)
{
	double body = 7 * 16 + 4 + 4 + 2 * 52 + (reverse ? 2 * 52 + 8 : 0);
	double contact = 8 + (double)R->substeps * (2 * 16 + 4 * 5 * 16);
	return body * R->nb + contact * R->nc;
}

static bool adj_launch(World& w, DiffState* D, const DiffStepRec* R, int iterations, hipStream_t s,
					   const float* stateIn, float* stateOut, int reverse, const double* seedIn, double* seedOut)
{
	int32_t MC = std::max(R->maxWorldPoints / 4, 1);
	int32_t MB = std::max(R->maxWorldBodies, 1);
	if (MC > kAdjMaxContacts || MB > kAdjMaxBodies)
	{
		set_error("a world has %d contacts and %d bodies; the parallel adjoint takes at most %d and %d", MC, MB,
				  kAdjMaxContacts, kAdjMaxBodies);
		return false;
	}
	AdjLayout L = adj_layout(MB, MC, R->substeps, iterations);
	if (!D->adjSlab.ensure((size_t)L.total * D->worlds) || !D->adjDSlab.ensure((size_t)L.dtotal * D->worlds))
		return false;
	RecView V = rec_view(D, R, iterations);
	w.mark_begin(reverse ? M3DX_K_ADJOINT : M3DX_K_ADJ_FORWARD, adj_bytes(R, reverse));
	if (reverse)
		k_adj_step<1><<<D->worlds, kAdjThreads, 0, s>>>(V, L, D->adjSlab.p, D->adjDSlab.p, stateIn, stateOut, seedIn,
														 seedOut, D->massG.p, D->fricG.p);
	else
		k_adj_step<0><<<D->worlds, kAdjThreads, 0, s>>>(V, L, D->adjSlab.p, D->adjDSlab.p, stateIn, stateOut, seedIn,
														 seedOut, D->massG.p, D->fricG.p);
	w.mark_end();
	return hip_ok(hipGetLastError(), "adjoint launch");
}

extern "C"
{

// Steps between the body-state checkpoints the backward pass restarts from
// (default 10); the states in between are recomputed. Returns 0 or -1.
int m3dx_diff_set_checkpoint(m3d_world* world, int every)
{
	World& w = world->w;
	if (diff_refused(w))
		return -1;
	if (w.diff == nullptr || every < 1)
	{
		set_error("no differentiable recording, or checkpoint interval below 1");
		return -1;
	}
	w.diff->checkpointEvery = every;
	return 0;
}

// Vector-Jacobian product of the recorded steps (the parallel adjoint), in
// double precision along the engine's float path: given the adjoint of the
// final state (host, 13 per body: x, q, v, w), returns the adjoint of the
// first recorded step's input state (13 per body; its v is the gradient with
// respect to the initial velocities) and the gradients with respect to each
// body's mass and friction (summed over the steps). Any pointer but seed may
// be null. Returns 0, or -1 (m3dx_last_error).
int m3dx_diff_backward_f64(m3d_world* world, const double* seed, double* stateAdj, double* massGrad,
						   double* frictionGrad)
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
	const int32_t n = D->steps[0]->nb;
	const int N = (int)D->steps.size();
	const int K = std::max(D->checkpointEvery, 1);
	const int iterations = iterations_of(w);
	for (const DiffStepRec* R : D->steps)
		if (R->nb != n)
		{
			set_error("body count changed while recording");
			return -1;
		}
	size_t st = (size_t)n * kOuts;
	if (!D->seedA.ensure(st) || !D->seedB.ensure(st) || !D->massG.ensure(std::max(n, 1)) ||
		!D->fricG.ensure(std::max(n, 1)) || !D->segStates.ensure(st * (size_t)K))
		return -1;
	double* sa = D->seedA.p;
	double* sb = D->seedB.p;
	hipMemcpyAsync(sa, seed, st * sizeof(double), hipMemcpyHostToDevice, s);
	hipMemsetAsync(D->massG.p, 0, (size_t)n * sizeof(double), s);
	hipMemsetAsync(D->fricG.p, 0, (size_t)n * sizeof(double), s);
	for (int seg = ((N - 1) / K) * K; seg >= 0; seg -= K)
	{
		int end = std::min(seg + K, N);
		// recompute the states inside the segment from its checkpoint
		for (int k = seg; k < end - 1; ++k)
			if (!adj_launch(w, D, D->steps[k], iterations, s, k == seg ? nullptr : D->segStates.p + st * (k - seg),
							D->segStates.p + st * (k - seg + 1), 0, nullptr, nullptr))
				return -1;
		for (int k = end - 1; k >= seg; --k)
		{
			if (!adj_launch(w, D, D->steps[k], iterations, s, k == seg ? nullptr : D->segStates.p + st * (k - seg),
							nullptr, 1, sa, sb))
				return -1;
			std::swap(sa, sb);
		}
	}
	std::vector<double> ha(st), hm((size_t)n), hf((size_t)n);
	hipMemcpyAsync(ha.data(), sa, st * sizeof(double), hipMemcpyDeviceToHost, s);
	hipMemcpyAsync(hm.data(), D->massG.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s);
	hipMemcpyAsync(hf.data(), D->fricG.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s);
	if (!hip_ok(hipStreamSynchronize(s), "adjoint"))
		return -1;
	w.collect_marks();
	if (stateAdj != nullptr)
		memcpy(stateAdj, ha.data(), st * sizeof(double));
	if (massGrad != nullptr)
		memcpy(massGrad, hm.data(), (size_t)n * sizeof(double));
	if (frictionGrad != nullptr)
		memcpy(frictionGrad, hf.data(), (size_t)n * sizeof(double));
	return 0;
}

// m3dx_diff_backward_f64 with float arrays.
int m3dx_diff_backward(m3d_world* world, const float* seed, float* stateAdj, float* massGrad, float* frictionGrad)
{
	World& w = world->w;
	if (diff_refused(w))
		return -1;
	if (w.diff == nullptr || w.diff->steps.empty())
	{
		set_error("no recorded steps");
		return -1;
	}
	const size_t n = (size_t)w.diff->steps[0]->nb;
	std::vector<double> sd(seed, seed + n * kOuts), a(n * kOuts), m(n), f(n);
	if (m3dx_diff_backward_f64(world, sd.data(), a.data(), m.data(), f.data()) != 0)
		return -1;
	for (size_t k = 0; k < n * kOuts; ++k)
		if (stateAdj != nullptr)
			stateAdj[k] = (float)a[k];
	for (size_t k = 0; k < n; ++k)
	{
		if (massGrad != nullptr)
			massGrad[k] = (float)m[k];
		if (frictionGrad != nullptr)
			frictionGrad[k] = (float)f[k];
	}
	return 0;
}

// The recorded steps replayed forward by the same kernels (float, the
// engine's arithmetic) from the state recorded at step `first`: the state
// after step first + count - 1 (host, 13 per body). Returns 0 or -1.
int m3dx_diff_forward(m3d_world* world, int first, int count, float* stateOut)
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
	const int32_t n = D->steps[first]->nb;
	size_t st = (size_t)n * kOuts;
	if (!D->segStates.ensure(2 * st) || !D->massG.ensure(std::max(n, 1)) || !D->fricG.ensure(std::max(n, 1)))
		return -1;
	float* a = D->segStates.p;
	float* b = a + st;
	for (int k = 0; k < count; ++k)
	{
		if (!adj_launch(w, D, D->steps[first + k], iterations_of(w), s, k == 0 ? nullptr : a, b, 0, nullptr, nullptr))
			return -1;
		std::swap(a, b);
	}
	if (!hip_ok(hipMemcpyAsync(stateOut, a, st * sizeof(float), hipMemcpyDeviceToHost, s), "forward output") ||
		!hip_ok(hipStreamSynchronize(s), "forward"))
		return -1;
	w.collect_marks();
	return 0;
}

} // extern "C"
