// Differentiable replay of the XPBD solve of one step (the four substeps of
// src/xpbd.cpp:898-974 for serial islands and :1167-1369 for colored ones:
// integrate, prepare, normal projection iterations, stiction, velocity
// reconstruction, velocity pass) over a fixed contact set and coloring: the
// solver contacts and manifolds the engine's narrowphase produced for that
// step, and the colors and packet forms its solve used.
//
// Everything is written once over a generic scalar T:
//   float      the same expressions in the same order as mx_solver.h, so a
//              replay reproduces the engine's solve bit for bit;
//   double     the FP64 replay used for finite differences;
//   Ad<F>      a reverse-mode tape: every arithmetic result is a node with
//              its two partial derivatives; walking the tape backwards
//              gives the adjoint of the whole step.
// Branch decisions (sep < 0, caps, friction cone, slip coherence) are taken
// on values, so the adjoint is that of the piecewise-smooth step along the
// branches the forward pass took.
#pragma once

#include "mx_common.h"

#include <math.h>

// Loops over a manifold's (at most four) points; unrolling them would let
// fixed-size point arrays live in registers, at a large compile-time cost for
// the taped instantiation.
#ifndef MX_UNROLL_POINTS
#define MX_UNROLL_POINTS
#endif

namespace mx
{
namespace dd
{

// ------------------------------------------------------------ tape

constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// One tape node: its two inputs and the partial derivatives (one 16-byte
// store for F = float).
template <class F>
struct Node
{
	uint32_t a, b;
	F da, db;
};

// Node k of one lane at nodes[k * stride]. The count is 64-bit so that, by
// type, no node store can alias it: with the header in LDS at one known
// address, the compiler keeps the count in a register across operations.
template <class F>
struct Tape
{
	Node<F>* nodes;
	uint64_t stride;
	uint64_t n;
	uint64_t cap;
	int overflow;
};

// Tape scalar: a value and the node that produced it (kNoNode: a constant).
template <class F>
struct Ad
{
	F v;
	uint32_t id;
};

// The replaying lane's tape header (one active lane per workgroup).
__device__ __forceinline__ Tape<float>& ad_tape()
{
	__shared__ Tape<float> tape;
	return tape;
}

template <class F>
__device__ __forceinline__ uint32_t push_node(uint32_t a, F da, uint32_t b, F db)
{
	Tape<F>& t = ad_tape();
	uint64_t n = t.n;
	if (n >= t.cap)
	{
		t.overflow = 1;
		return kNoNode;
	}
	Node<F> e;
	e.a = a;
	e.b = b;
	e.da = da;
	e.db = db;
	t.nodes[n * t.stride] = e;
	t.n = n + 1;
	return (uint32_t)n;
}

template <class F>
__device__ __forceinline__ Ad<F> ad_make(F v, uint32_t a, F da, uint32_t b, F db)
{
	Ad<F> r;
	r.v = v;
	r.id = (a == kNoNode && b == kNoNode) ? kNoNode : push_node<F>(a, da, b, db);
	return r;
}

template <class F>
MX_HD Ad<F> ad_const(F v)
{
	Ad<F> r;
	r.v = v;
	r.id = kNoNode;
	return r;
}

// a leaf (parameter or state input): a node without inputs
template <class F>
__device__ __forceinline__ Ad<F> ad_leaf(F v)
{
	Ad<F> r;
	r.v = v;
	r.id = push_node<F>(kNoNode, (F)0, kNoNode, (F)0);
	return r;
}

template <class F>
__device__ __forceinline__ Ad<F> operator+(Ad<F> x, Ad<F> y)
{
	return ad_make<F>(x.v + y.v, x.id, (F)1, y.id, (F)1);
}
template <class F>
__device__ __forceinline__ Ad<F> operator-(Ad<F> x, Ad<F> y)
{
	return ad_make<F>(x.v - y.v, x.id, (F)1, y.id, (F)-1);
}
template <class F>
__device__ __forceinline__ Ad<F> operator*(Ad<F> x, Ad<F> y)
{
	return ad_make<F>(x.v * y.v, x.id, y.v, y.id, x.v);
}
template <class F>
__device__ __forceinline__ Ad<F> operator/(Ad<F> x, Ad<F> y)
{
	F r = x.v / y.v;
	return ad_make<F>(r, x.id, (F)1 / y.v, y.id, -r / y.v);
}
template <class F>
__device__ __forceinline__ Ad<F> operator-(Ad<F> x)
{
	return ad_make<F>(-x.v, x.id, (F)-1, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator+(Ad<F> x, F c)
{
	return ad_make<F>(x.v + c, x.id, (F)1, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator+(F c, Ad<F> x)
{
	return ad_make<F>(c + x.v, x.id, (F)1, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator-(Ad<F> x, F c)
{
	return ad_make<F>(x.v - c, x.id, (F)1, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator-(F c, Ad<F> x)
{
	return ad_make<F>(c - x.v, x.id, (F)-1, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator*(Ad<F> x, F c)
{
	return ad_make<F>(x.v * c, x.id, c, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator*(F c, Ad<F> x)
{
	return ad_make<F>(c * x.v, x.id, c, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator/(Ad<F> x, F c)
{
	return ad_make<F>(x.v / c, x.id, (F)1 / c, kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> operator/(F c, Ad<F> x)
{
	F r = c / x.v;
	return ad_make<F>(r, x.id, -r / x.v, kNoNode, (F)0);
}

// value of a scalar of any kind
MX_HD float val(float x) { return x; }
MX_HD double val(double x) { return x; }
template <class F>
MX_HD F val(const Ad<F>& x)
{
	return x.v;
}

// constants of the scalar's precision
template <class T>
struct Num;
template <>
struct Num<float>
{
	typedef float F;
	static MX_HD float c(double x) { return (float)x; }
};
template <>
struct Num<double>
{
	typedef double F;
	static MX_HD double c(double x) { return x; }
};
template <class G>
struct Num<Ad<G>>
{
	typedef G F;
	static MX_HD Ad<G> c(double x) { return ad_const<G>((G)x); }
};

// elementary functions; float keeps the engine's correctly rounded forms
MX_HD float sqrt_(float x) { return sqrtf(x); }
MX_HD double sqrt_(double x) { return sqrt(x); }
template <class F>
__device__ __forceinline__ Ad<F> sqrt_(Ad<F> x)
{
	F r = sqrt_(x.v);
	return ad_make<F>(r, x.id, r > (F)0 ? (F)0.5 / r : (F)0, kNoNode, (F)0);
}
MX_HD float sin_(float x) { return (float)sin((double)x); }
MX_HD double sin_(double x) { return sin(x); }
MX_HD float cos_(float x) { return (float)cos((double)x); }
MX_HD double cos_(double x) { return cos(x); }
template <class F>
__device__ __forceinline__ Ad<F> sin_(Ad<F> x)
{
	return ad_make<F>(sin_(x.v), x.id, cos_(x.v), kNoNode, (F)0);
}
template <class F>
__device__ __forceinline__ Ad<F> cos_(Ad<F> x)
{
	return ad_make<F>(cos_(x.v), x.id, -sin_(x.v), kNoNode, (F)0);
}
MX_HD float atan2_(float y, float x) { return (float)atan2((double)y, (double)x); }
MX_HD double atan2_(double y, double x) { return atan2(y, x); }
template <class F>
__device__ __forceinline__ Ad<F> atan2_(Ad<F> y, Ad<F> x)
{
	F r2 = x.v * x.v + y.v * y.v;
	F dy = r2 > (F)0 ? x.v / r2 : (F)0;
	F dx = r2 > (F)0 ? -y.v / r2 : (F)0;
	return ad_make<F>(atan2_(y.v, x.v), y.id, dy, x.id, dx);
}
MX_HD float fabs_(float x) { return fabsf(x); }
MX_HD double fabs_(double x) { return fabs(x); }
template <class F>
__device__ __forceinline__ Ad<F> fabs_(Ad<F> x)
{
	return ad_make<F>(x.v < (F)0 ? -x.v : x.v, x.id, x.v < (F)0 ? (F)-1 : (F)1, kNoNode, (F)0);
}

// ------------------------------------------------------------ forward tangents

// Forward-mode scalar: a value and its derivative along one input
// direction. The value part is computed by exactly the plain expressions of
// its type F, so a Dual replay takes the same branches as the plain one; the
// derivative may be carried in a wider type G (float values with double
// derivatives differentiate the float path in double precision).
template <class F, class G = F>
struct Dual
{
	F v;
	G d;
};

template <class F, class G = F>
MX_HD Dual<F, G> dual(F v, G d)
{
	Dual<F, G> r;
	r.v = v;
	r.d = d;
	return r;
}

template <class F, class G>
MX_HD Dual<F, G> operator+(Dual<F, G> x, Dual<F, G> y)
{
	return dual<F, G>(x.v + y.v, x.d + y.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator-(Dual<F, G> x, Dual<F, G> y)
{
	return dual<F, G>(x.v - y.v, x.d - y.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator*(Dual<F, G> x, Dual<F, G> y)
{
	return dual<F, G>(x.v * y.v, x.d * (G)y.v + (G)x.v * y.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator/(Dual<F, G> x, Dual<F, G> y)
{
	F r = x.v / y.v;
	return dual<F, G>(r, (x.d - (G)r * y.d) / (G)y.v);
}
template <class F, class G>
MX_HD Dual<F, G> operator-(Dual<F, G> x)
{
	return dual<F, G>(-x.v, -x.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator+(Dual<F, G> x, F c)
{
	return dual<F, G>(x.v + c, x.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator+(F c, Dual<F, G> x)
{
	return dual<F, G>(c + x.v, x.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator-(Dual<F, G> x, F c)
{
	return dual<F, G>(x.v - c, x.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator-(F c, Dual<F, G> x)
{
	return dual<F, G>(c - x.v, -x.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator*(Dual<F, G> x, F c)
{
	return dual<F, G>(x.v * c, x.d * (G)c);
}
template <class F, class G>
MX_HD Dual<F, G> operator*(F c, Dual<F, G> x)
{
	return dual<F, G>(c * x.v, (G)c * x.d);
}
template <class F, class G>
MX_HD Dual<F, G> operator/(Dual<F, G> x, F c)
{
	return dual<F, G>(x.v / c, x.d / (G)c);
}
template <class F, class G>
MX_HD Dual<F, G> operator/(F c, Dual<F, G> x)
{
	F r = c / x.v;
	return dual<F, G>(r, -(G)r * x.d / (G)x.v);
}

template <class F, class G>
MX_HD F val(const Dual<F, G>& x)
{
	return x.v;
}

template <class P, class G>
struct Num<Dual<P, G>>
{
	typedef P F;
	static MX_HD Dual<P, G> c(double x) { return dual<P, G>((P)x, (G)0); }
};

template <class F, class G>
MX_HD Dual<F, G> sqrt_(Dual<F, G> x)
{
	F r = sqrt_(x.v);
	return dual<F, G>(r, r > (F)0 ? x.d * ((G)0.5 / (G)r) : (G)0);
}
template <class F, class G>
MX_HD Dual<F, G> sin_(Dual<F, G> x)
{
	return dual<F, G>(sin_(x.v), cos_((G)x.v) * x.d);
}
template <class F, class G>
MX_HD Dual<F, G> cos_(Dual<F, G> x)
{
	return dual<F, G>(cos_(x.v), -sin_((G)x.v) * x.d);
}
template <class F, class G>
MX_HD Dual<F, G> atan2_(Dual<F, G> y, Dual<F, G> x)
{
	G xv = (G)x.v, yv = (G)y.v;
	G r2 = xv * xv + yv * yv;
	G d = r2 > (G)0 ? (xv * y.d - yv * x.d) / r2 : (G)0;
	return dual<F, G>(atan2_(y.v, x.v), d);
}
template <class F, class G>
MX_HD Dual<F, G> fabs_(Dual<F, G> x)
{
	return x.v < (F)0 ? dual<F, G>(-x.v, -x.d) : x;
}

// ------------------------------------------------------------ vectors

template <class T>
struct V3
{
	T x, y, z;
};
template <class T>
struct Q4
{
	T x, y, z, w;
};
template <class T>
struct M3
{
	V3<T> cx, cy, cz;
};

template <class T>
MX_HD V3<T> v3(T x, T y, T z)
{
	V3<T> r = { x, y, z };
	return r;
}
template <class T>
MX_HD V3<T> add(V3<T> a, V3<T> b)
{
	return v3<T>(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class T>
MX_HD V3<T> sub(V3<T> a, V3<T> b)
{
	return v3<T>(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class T>
MX_HD V3<T> neg(V3<T> a)
{
	return v3<T>(-a.x, -a.y, -a.z);
}
template <class T, class S>
MX_HD V3<T> scale(S s, V3<T> v)
{
	return v3<T>(s * v.x, s * v.y, s * v.z);
}
template <class T, class S>
MX_HD V3<T> mul_add(V3<T> a, S s, V3<T> b)
{
	return v3<T>(a.x + s * b.x, a.y + s * b.y, a.z + s * b.z);
}
template <class T>
MX_HD T dot(V3<T> a, V3<T> b)
{
	return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
MX_HD V3<T> cross(V3<T> a, V3<T> b)
{
	return v3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T>
MX_HD T length_sq(V3<T> v)
{
	return dot(v, v);
}
template <class T>
MX_HD T length(V3<T> v)
{
	return sqrt_(dot(v, v));
}
template <class T>
MX_HD Q4<T> quat_mul(Q4<T> a, Q4<T> b)
{
	Q4<T> r;
	r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
	r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
	r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
	r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
	return r;
}
template <class T>
MX_HD Q4<T> quat_conj(Q4<T> q)
{
	Q4<T> r = { -q.x, -q.y, -q.z, q.w };
	return r;
}
template <class T>
MX_HD V3<T> rotate(Q4<T> q, V3<T> v)
{
	typedef typename Num<T>::F F;
	V3<T> u = v3<T>(q.x, q.y, q.z);
	V3<T> t = scale((F)2, cross(u, v));
	return add(add(v, scale(q.w, t)), cross(u, t));
}
template <class T>
MX_HD M3<T> quat_to_mat3(Q4<T> q)
{
	typedef typename Num<T>::F F;
	T x = q.x, y = q.y, z = q.z, w = q.w;
	T x2 = x + x, y2 = y + y, z2 = z + z;
	T xx = x * x2, xy = x * y2, xz = x * z2;
	T yy = y * y2, yz = y * z2, zz = z * z2;
	T wx = w * x2, wy = w * y2, wz = w * z2;
	M3<T> m;
	m.cx = v3<T>((F)1 - (yy + zz), xy + wz, xz - wy);
	m.cy = v3<T>(xy - wz, (F)1 - (xx + zz), yz + wx);
	m.cz = v3<T>(xz + wy, yz - wx, (F)1 - (xx + yy));
	return m;
}
template <class T>
MX_HD V3<T> mat3_mulv(const M3<T>& m, V3<T> v)
{
	return add(add(scale(v.x, m.cx), scale(v.y, m.cy)), scale(v.z, m.cz));
}

// src/xpbd.cpp:21-25
template <class T>
MX_HD Q4<T> renorm_fast(Q4<T> q)
{
	typedef typename Num<T>::F F;
	T s = (F)1.5 - (F)0.5 * (q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
	Q4<T> r = { q.x * s, q.y * s, q.z * s, q.w * s };
	return r;
}

template <class T>
MX_HD Q4<T> rotate_by(Q4<T> q, V3<T> dw)
{
	typedef typename Num<T>::F F;
	Q4<T> wq = { dw.x, dw.y, dw.z, Num<T>::c(0.0) };
	Q4<T> dq = quat_mul(wq, q);
	Q4<T> r = { q.x + (F)0.5 * dq.x, q.y + (F)0.5 * dq.y, q.z + (F)0.5 * dq.z, q.w + (F)0.5 * dq.w };
	return renorm_fast(r);
}

// I^-1_world = R diag(d) R^T (src/m3d_internal.h:126-146)
template <class T>
MX_HD M3<T> inv_inertia_world(Q4<T> rot, V3<T> d)
{
	M3<T> r = quat_to_mat3(rot);
	M3<T> rd;
	rd.cx = scale(d.x, r.cx);
	rd.cy = scale(d.y, r.cy);
	rd.cz = scale(d.z, r.cz);
	M3<T> m;
	m.cx = v3<T>(rd.cx.x * r.cx.x + rd.cy.x * r.cy.x + rd.cz.x * r.cz.x,
				 rd.cx.y * r.cx.x + rd.cy.y * r.cy.x + rd.cz.y * r.cz.x,
				 rd.cx.z * r.cx.x + rd.cy.z * r.cy.x + rd.cz.z * r.cz.x);
	m.cy = v3<T>(rd.cx.x * r.cx.y + rd.cy.x * r.cy.y + rd.cz.x * r.cz.y,
				 rd.cx.y * r.cx.y + rd.cy.y * r.cy.y + rd.cz.y * r.cz.y,
				 rd.cx.z * r.cx.y + rd.cy.z * r.cy.y + rd.cz.z * r.cz.y);
	m.cz = v3<T>(rd.cx.x * r.cx.z + rd.cy.x * r.cy.z + rd.cz.x * r.cz.z,
				 rd.cx.y * r.cx.z + rd.cy.y * r.cy.z + rd.cz.y * r.cz.z,
				 rd.cx.z * r.cx.z + rd.cy.z * r.cy.z + rd.cz.z * r.cz.z);
	return m;
}

// ------------------------------------------------------------ step data

// Inputs of one world's step (plain floats, as recorded; device or host
// memory). Bodies: the world's range; contacts: its solver contacts in
// ascending engine order (the serial order of src/world.cpp:2012-2058 per
// island; islands touch disjoint dynamic bodies, so one ascending pass
// equals island-by-island).
struct StepIn
{
	int32_t nb;           // bodies
	const float4* pos;    // xyz, invMass
	const float4* rot;
	const float4* vel;    // xyz, gravityScale
	const float4* angVel; // xyz, mass
	const float4* invIL;  // xyz, linearDamping
	const float* angDamp;
	const float4* shape;  // halfExtents xyz, radius
	const float4* bmat;   // halfHeight, friction, restitution, rollingResistance
	const uint32_t* flags; // BF_* bits
	int32_t nc;           // solver contacts
	const int2* bodies;   // local body indices, a < b
	const float4* normal; // xyz, pointCount bits
	const float4* cmat;   // friction, restitution, rolling, CF_* bits
	// per contact, 4 point slots each, [c * 4 + j]
	const float4* pA;     // localAnchorA, separationOffset
	const float4* pB;     // localAnchorB
	const float4* pFA;    // friction anchors at step start
	const float4* pFB;
	const float4* pSlip;  // slip direction
	// per contact: the schedule word (normal projection form, pass-order
	// key; sched_form, sched_key); null: every contact serial, key 0
	const int32_t* sched;
	// the manifold fields above and the schedule words hold one block per
	// substep (the manifold each substep's prepare left, re-collided or not):
	// substep s reads the block at s * cStride contacts
	int64_t cStride;
	float dt;
	int substeps;
	m3d_vec3 gravity;
};

// Mutable replay state of one world (scratch of the replaying lane).
template <class T>
struct StepState
{
	V3<T>* x;
	Q4<T>* q;
	V3<T>* v;
	V3<T>* w;
	V3<T>* px; // previous substep
	Q4<T>* pq;
	V3<T>* pv;
	V3<T>* pw;
	M3<T>* iw;
	T* invMass;
	V3<T>* il;   // inverse local inertia
	T* mu;       // per contact: combined friction
	// per point [c * 4 + j]
	T* lambdaN;
	T* lambdaT;
	V3<T>* rA;
	V3<T>* rB;
	V3<T>* jA;
	V3<T>* jB;
	T* wN;
	float4* fA; // friction anchors (no derivative: re-anchoring copies constants)
	float4* fB;
	float4* slip;
};

MX_HD V3<float> f3(float4 v) { return v3<float>(v.x, v.y, v.z); }

template <class T>
MX_HD V3<T> lift(float4 v)
{
	return v3<T>(Num<T>::c(v.x), Num<T>::c(v.y), Num<T>::c(v.z));
}

MX_HD bool is_dynamic(const StepIn& in, int32_t i) { return bf_type(in.flags[i]) == M3D_BODY_DYNAMIC; }

// ------------------------------------------------------------ stages
//
// Each stage is written once on values: one body's or one contact's inputs
// and outputs in registers. The array-based replay below (solve_step) loads
// them from a world's state and stores them back; the per-contact adjoint
// evaluates the same functions on forward tangents.

// IntegrateBodySubstep (src/xpbd.cpp:587-705) for a body without forces,
// torques or gyroscopic term (isotropic inverse inertia); mx_solver.h
// integrate_body. The previous-substep copies are the inputs themselves.
template <class T>
MX_HD void integrate_op(V3<T>& x, Q4<T>& q, V3<T>& v, V3<T>& w, V3<T> il, M3<T>& iw, float h, float gravityScale,
						float linDamp, float angDamp, m3d_vec3 gravity)
{
	typedef typename Num<T>::F F;
	V3<T> vn = mul_add(v, (F)(h * gravityScale), v3<T>(Num<T>::c(gravity.x), Num<T>::c(gravity.y),
														Num<T>::c(gravity.z)));
	vn = scale((F)(1.0f / (1.0f + h * linDamp)), vn);
	V3<T> wn = scale((F)(1.0f / (1.0f + h * angDamp)), w);
	V3<T> xn = mul_add(x, (F)h, vn);
	Q4<T> rot = q;
	T wlen = length(wn);
	if (val(wlen) > (F)1e-9)
	{
		T half = (F)(0.5f * h) * wlen;
		T sh, ch;
		if (val(half) < (F)0.02)
		{
			T h2 = half * half;
			sh = half * ((F)1 - h2 * (F)(1.0f / 6.0f));
			ch = (F)1 - h2 * (F)0.5;
		}
		else
		{
			sh = sin_(half);
			ch = cos_(half);
		}
		T s = sh / wlen;
		Q4<T> dq = { wn.x * s, wn.y * s, wn.z * s, ch };
		rot = renorm_fast(quat_mul(dq, rot));
	}
	iw = inv_inertia_world(rot, il);
	x = xn;
	q = rot;
	v = vn;
	w = wn;
}

// contact arm of src/xpbd.cpp:100-116 (mx_solver.h contact_arm)
template <class T>
MX_HD V3<T> contact_arm(int shapeType, float radius, float halfHeight, Q4<T> rot, float4 local, V3<T> n, float sign)
{
	typedef typename Num<T>::F F;
	if (shapeType == M3D_SHAPE_SPHERE)
		return scale((F)(sign * radius), n);
	if (shapeType == M3D_SHAPE_CAPSULE)
	{
		float ay = local.y < -halfHeight ? -halfHeight : (local.y > halfHeight ? halfHeight : local.y);
		V3<T> axis = rotate(rot, v3<T>(Num<T>::c(0.0), Num<T>::c(1.0), Num<T>::c(0.0)));
		return add(scale((F)ay, axis), scale((F)(sign * radius), n));
	}
	return rotate(rot, lift<T>(local));
}

template <class T>
struct Body
{
	V3<T> p;
	Q4<T> q;
	V3<T> v;
	V3<T> w;
	T invMass;
	M3<T> iw;
};

// Shape and manifold data of one contact in one substep (plain floats).
struct ContactGeo
{
	float4 n;     // normal xyz, point count bits
	float4 mat;   // friction, restitution, rolling, CF_* bits
	// point j's localAnchorA and separationOffset, localAnchorB: lane j of
	// vector registers (a run-time j does not put the geometry in scratch)
	mx_float4v pax, pay, paz, paw, pbx, pby, pbz, pbw;
	MX_HD float4 pa(int j) const { return make_float4(pax[j], pay[j], paz[j], paw[j]); }
	MX_HD float4 pb(int j) const { return make_float4(pbx[j], pby[j], pbz[j], pbw[j]); }
	float radiusA, radiusB, hhA, hhB;
	int tA, tB;
	int count;
	bool dynA, dynB;
};

MX_HD ContactGeo contact_geo(const StepIn& in, int32_t c)
{
	ContactGeo g;
	int2 ab = in.bodies[c];
	g.n = in.normal[c];
	g.mat = in.cmat[c];
	g.count = f2i(g.n.w);
	for (int j = 0; j < 4; ++j)
	{
		const float4 a = in.pA[c * 4 + j], b = in.pB[c * 4 + j];
		g.pax[j] = a.x;
		g.pay[j] = a.y;
		g.paz[j] = a.z;
		g.paw[j] = a.w;
		g.pbx[j] = b.x;
		g.pby[j] = b.y;
		g.pbz[j] = b.z;
		g.pbw[j] = b.w;
	}
	g.radiusA = in.shape[ab.x].w;
	g.radiusB = in.shape[ab.y].w;
	g.hhA = in.bmat[ab.x].x;
	g.hhB = in.bmat[ab.y].x;
	g.tA = bf_shape(in.flags[ab.x]);
	g.tB = bf_shape(in.flags[ab.y]);
	g.dynA = is_dynamic(in, ab.x);
	g.dynB = is_dynamic(in, ab.y);
	return g;
}

// Frozen arms, Jacobians and normal mass of one point (src/xpbd.cpp:845-861)
template <class T>
struct PointPrep
{
	V3<T> rA, rB, jA, jB;
	T wN;
};

template <class T>
MX_HD PointPrep<T> prepare_point(const ContactGeo& g, int j, Q4<T> qa, Q4<T> qb, const M3<T>& iwa, const M3<T>& iwb,
								 T imA, T imB)
{
	V3<T> n = lift<T>(g.n);
	PointPrep<T> r;
	r.rA = contact_arm(g.tA, g.radiusA, g.hhA, qa, g.pa(j), n, 1.0f);
	r.rB = contact_arm(g.tB, g.radiusB, g.hhB, qb, g.pb(j), n, -1.0f);
	V3<T> crA = cross(r.rA, n);
	V3<T> crB = cross(r.rB, n);
	r.jA = mat3_mulv(iwa, crA);
	r.jB = mat3_mulv(iwb, crB);
	r.wN = imA + imB + dot(crA, r.jA) + dot(crB, r.jB);
	return r;
}

// src/xpbd.cpp:45-58
template <class T>
MX_HD void axis_correction(V3<T>& pos, Q4<T>& rot, T invMass, T dl, V3<T> n, V3<T> angJ)
{
	typedef typename Num<T>::F F;
	if (val(invMass) == (F)0)
		return;
	pos = mul_add(pos, dl * invMass, n);
	rot = rotate_by(rot, scale(dl, angJ));
}

// SolveContactNormals, serial form (src/xpbd.cpp:122-163): the contact's
// points in order against the two bodies' positions and rotations.
template <class T>
MX_HD void normals_op(const ContactGeo& g, V3<T>& pa, Q4<T>& qa, V3<T>& pb, Q4<T>& qb, T imA, T imB, T* lambdaN,
					  const T* wN, const V3<T>* jA, const V3<T>* jB, float maxCorrection)
{
	typedef typename Num<T>::F F;
	V3<T> n = lift<T>(g.n);
MX_UNROLL_POINTS
	for (int j = 0; j < 4; ++j)
	{
		if (j >= g.count)
			break;
		T w = wN[j];
		if (!(val(w) > (F)0))
			continue;
		float4 a4 = g.pa(j);
		V3<T> rA = rotate(qa, lift<T>(a4));
		V3<T> rB = rotate(qb, lift<T>(g.pb(j)));
		T sep = Num<T>::c(a4.w) + dot(sub(add(pb, rB), add(pa, rA)), n);
		if (!(val(sep) < (F)0))
			continue;
		if (val(sep) < (F)(-maxCorrection))
			sep = Num<T>::c(-maxCorrection);
		T dl = -sep / w;
		lambdaN[j] = lambdaN[j] + dl;
		axis_correction(pb, qb, imB, dl, n, jB[j]);
		axis_correction(pa, qa, imA, -dl, n, jA[j]);
	}
}

// SolvePacketNormals (src/xpbd.cpp:1047-1141) seen from one lane: the
// contacts of one color are solved side by side in packets of 8, each lane
// taking its points in sequence up to the packet's largest point count
// (maxPoints). A point beyond the lane's own count is the zero point with
// invW = 1 and a zero gate: it moves nothing, yet both rotations are still
// renormalized. Both bodies' lane copies are updated (a static body's copy
// only by the renormalization; the caller keeps the dynamic ones). The
// expressions are those of mx_solver.h solve_normals_packet_in, in order.
template <class T>
MX_HD void normals_packet_op(const ContactGeo& g, int maxPoints, V3<T>& pa, Q4<T>& qa, V3<T>& pb, Q4<T>& qb, T imA,
							 T imB, T* lambdaN, const T* wN, const V3<T>* jA, const V3<T>* jB, float maxCorrection)
{
	typedef typename Num<T>::F F;
	const T zero = Num<T>::c(0.0);
	const V3<T> zero3 = v3<T>(zero, zero, zero);
	const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	const F nx = (F)g.n.x, ny = (F)g.n.y, nz = (F)g.n.z;
MX_UNROLL_POINTS
	for (int j = 0; j < 4; ++j)
	{
		if (j >= maxPoints)
			break;
		const bool on = j < g.count;
		const float4 a4 = on ? g.pa(j) : zero4;
		const float4 b4 = on ? g.pb(j) : zero4;
		const V3<T> ja = on ? jA[j] : zero3;
		const V3<T> jb = on ? jB[j] : zero3;
		const F active = on ? (F)1 : (F)0;
		const T invW = (on && val(wN[j]) > (F)0) ? (F)1 / wN[j] : Num<T>::c(1.0);

		F vx = (F)a4.x, vy = (F)a4.y, vz = (F)a4.z;
		T tx = (F)2 * (qa.y * vz - qa.z * vy);
		T ty = (F)2 * (qa.z * vx - qa.x * vz);
		T tz = (F)2 * (qa.x * vy - qa.y * vx);
		T rax = vx + qa.w * tx + (qa.y * tz - qa.z * ty);
		T ray = vy + qa.w * ty + (qa.z * tx - qa.x * tz);
		T raz = vz + qa.w * tz + (qa.x * ty - qa.y * tx);
		vx = (F)b4.x;
		vy = (F)b4.y;
		vz = (F)b4.z;
		tx = (F)2 * (qb.y * vz - qb.z * vy);
		ty = (F)2 * (qb.z * vx - qb.x * vz);
		tz = (F)2 * (qb.x * vy - qb.y * vx);
		T rbx = vx + qb.w * tx + (qb.y * tz - qb.z * ty);
		T rby = vy + qb.w * ty + (qb.z * tx - qb.x * tz);
		T rbz = vz + qb.w * tz + (qb.x * ty - qb.y * tx);

		T sep = (F)a4.w + (pb.x + rbx - pa.x - rax) * nx + (pb.y + rby - pa.y - ray) * ny +
				(pb.z + rbz - pa.z - raz) * nz;
		const F gate = val(sep) < (F)0 ? active : (F)0;
		T capped = val(sep) < (F)(-maxCorrection) ? Num<T>::c(-maxCorrection) : sep;
		T dl = -capped * invW * gate;
		if (on)
			lambdaN[j] = lambdaN[j] + dl;

		pb.x = pb.x + dl * imB * nx;
		pb.y = pb.y + dl * imB * ny;
		pb.z = pb.z + dl * imB * nz;
		pa.x = pa.x - dl * imA * nx;
		pa.y = pa.y - dl * imA * ny;
		pa.z = pa.z - dl * imA * nz;

		T dwx = dl * jb.x, dwy = dl * jb.y, dwz = dl * jb.z;
		T dqx = dwx * qb.w + dwy * qb.z - dwz * qb.y;
		T dqy = -dwx * qb.z + dwy * qb.w + dwz * qb.x;
		T dqz = dwx * qb.y - dwy * qb.x + dwz * qb.w;
		T dqw = -(dwx * qb.x + dwy * qb.y + dwz * qb.z);
		T x = qb.x + (F)0.5 * dqx, y = qb.y + (F)0.5 * dqy, z = qb.z + (F)0.5 * dqz, w = qb.w + (F)0.5 * dqw;
		T s = (F)1.5 - (F)0.5 * (x * x + y * y + z * z + w * w);
		qb.x = x * s;
		qb.y = y * s;
		qb.z = z * s;
		qb.w = w * s;

		dwx = -dl * ja.x;
		dwy = -dl * ja.y;
		dwz = -dl * ja.z;
		dqx = dwx * qa.w + dwy * qa.z - dwz * qa.y;
		dqy = -dwx * qa.z + dwy * qa.w + dwz * qa.x;
		dqz = dwx * qa.y - dwy * qa.x + dwz * qa.w;
		dqw = -(dwx * qa.x + dwy * qa.y + dwz * qa.z);
		x = qa.x + (F)0.5 * dqx;
		y = qa.y + (F)0.5 * dqy;
		z = qa.z + (F)0.5 * dqz;
		w = qa.w + (F)0.5 * dqw;
		s = (F)1.5 - (F)0.5 * (x * x + y * y + z * z + w * w);
		qa.x = x * s;
		qa.y = y * s;
		qa.z = z * s;
		qa.w = w * s;
	}
}

// A contact's schedule word in a step's record (one per contact and
// substep): bits 0-7 the normal projection's form, kFormSerial for the
// serial form or the packet's largest point count (0-4) for the packet form;
// bits 8-11 the pass-order key: the contact's color (0-7, 8 for the overflow
// pass) in a colored island, 0 in any other island.
constexpr int32_t kFormSerial = 0xFF;
constexpr int kOrderKeys = 9;
MX_HD int sched_form(const StepIn& in, int32_t c) { return in.sched != nullptr ? (in.sched[c] & 0xFF) : kFormSerial; }
MX_HD int sched_key(const StepIn& in, int32_t c) { return in.sched != nullptr ? ((in.sched[c] >> 8) & 0xF) : 0; }

// One more than the largest pass-order key of the step's contacts: the passes
// run key by key, ascending contacts within a key. That is every island's own
// order (src/xpbd.cpp:898-974 serial; :1269-1368 colored: colors in turn, then
// the overflow contacts), since islands touch disjoint dynamic bodies.
MX_HD int sched_keys(const StepIn& in)
{
	int k = 0;
	for (int32_t c = 0; c < in.nc; ++c)
	{
		int kc = sched_key(in, c) + 1;
		k = kc > k ? kc : k;
	}
	return k;
}

// The normal projection of contact c in its form.
template <class T>
MX_HD void normals_any(const StepIn& in, int32_t c, const ContactGeo& g, V3<T>& pa, Q4<T>& qa, V3<T>& pb, Q4<T>& qb,
					   T imA, T imB, T* lambdaN, const T* wN, const V3<T>* jA, const V3<T>* jB, float maxCorrection)
{
	const int form = sched_form(in, c);
	if (form == kFormSerial)
		normals_op(g, pa, qa, pb, qb, imA, imB, lambdaN, wN, jA, jB, maxCorrection);
	else
		normals_packet_op(g, form, pa, qa, pb, qb, imA, imB, lambdaN, wN, jA, jB, maxCorrection);
}

// src/xpbd.cpp:85-89
template <class T>
MX_HD T generalized_inv_mass(const Body<T>& b, V3<T> r, V3<T> axis)
{
	V3<T> rn = cross(r, axis);
	return b.invMass + dot(rn, mat3_mulv(b.iw, rn));
}

// src/xpbd.cpp:28-41
template <class T>
MX_HD void position_correction(Body<T>& b, V3<T> p, V3<T> r)
{
	typedef typename Num<T>::F F;
	if (val(b.invMass) == (F)0)
		return;
	b.p = mul_add(b.p, b.invMass, p);
	V3<T> dw = mat3_mulv(b.iw, cross(r, p));
	b.q = rotate_by(b.q, dw);
}

// ApplyContactFriction (src/xpbd.cpp:172-269), box-box pairs only; rA0 and
// rB0 are the frozen arms of point 0 (the speed gate), fA/fB the friction
// anchors (re-anchoring replaces them by the contact anchors).
template <class T>
MX_HD void friction_op(const ContactGeo& g, Body<T>& a, Body<T>& b, T mu, const T* lambdaN, T* lambdaT, V3<T> rA0,
					   V3<T> rB0, float4* fA, float4* fB)
{
	typedef typename Num<T>::F F;
	if (g.count == 0)
		return;
	V3<T> n = lift<T>(g.n);
	{
		const F maxSpeed2 = (F)(0.1f * 0.1f);
		V3<T> vrel = sub(add(b.v, cross(b.w, rB0)), add(a.v, cross(a.w, rA0)));
		V3<T> vt = sub(vrel, scale(dot(vrel, n), n));
		if (val(length_sq(vt)) > maxSpeed2)
		{
MX_UNROLL_POINTS
			for (int j = 0; j < 4; ++j)
			{
				if (j >= g.count)
					break;
				fA[j] = g.pa(j);
				fB[j] = g.pb(j);
			}
			return;
		}
	}
MX_UNROLL_POINTS
	for (int j = 0; j < 4; ++j)
	{
		if (j >= g.count)
			break;
		T lN = lambdaN[j];
		if (val(lN) <= (F)0)
			continue;
		V3<T> rA = rotate(a.q, lift<T>(fA[j]));
		V3<T> rB = rotate(b.q, lift<T>(fB[j]));
		V3<T> gap = sub(add(b.p, rB), add(a.p, rA));
		V3<T> dpt = sub(gap, scale(dot(gap, n), n));
		T lt = length(dpt);
		if (val(lt) <= (F)1e-9)
			continue;
		V3<T> t = scale((F)1 / lt, dpt);
		T wt = generalized_inv_mass(a, rA, t) + generalized_inv_mass(b, rB, t);
		if (val(wt) <= (F)0)
			continue;
		T needed = (F)kFrictionRelaxation * (lt / wt);
		T lT = lambdaT[j];
		T allowed = mu * lN - lT;
		if (val(allowed) <= (F)0)
			continue;
		T dL = val(needed) < val(allowed) ? needed : allowed;
		lambdaT[j] = lT + dL;
		V3<T> pt = scale(dL, t);
		position_correction(b, neg(pt), rB);
		position_correction(a, pt, rA);
		if (val(needed) > (F)2 * val(allowed))
		{
			fA[j] = g.pa(j);
			fB[j] = g.pb(j);
		}
	}
}

// ReconstructBodyVelocity (src/xpbd.cpp:865-892)
template <class T>
MX_HD void reconstruct_op(V3<T> x, Q4<T> q, V3<T> px, Q4<T> pq, float invH, V3<T>& v, V3<T>& w)
{
	typedef typename Num<T>::F F;
	v = scale((F)invH, sub(x, px));
	Q4<T> dq = quat_mul(q, quat_conj(pq));
	F sign = val(dq.w) >= (F)0 ? (F)1 : (F)-1;
	T v2 = dq.x * dq.x + dq.y * dq.y + dq.z * dq.z;
	if (val(v2) > (F)4e-4)
	{
		T vlen = sqrt_(v2);
		T angle = (F)2 * atan2_(vlen, fabs_(dq.w));
		w = scale(sign * (F)invH * angle / vlen, v3<T>(dq.x, dq.y, dq.z));
	}
	else
	{
		T corr = (F)1 + v2 * (F)(1.0f / 6.0f);
		w = scale((F)2 * (F)invH * sign * corr, v3<T>(dq.x, dq.y, dq.z));
	}
}

// src/xpbd.cpp:61-82
template <class T>
MX_HD void axis_velocity(Body<T>& b, T j, V3<T> n, V3<T> angJ)
{
	typedef typename Num<T>::F F;
	if (val(b.invMass) == (F)0)
		return;
	b.v = mul_add(b.v, j * b.invMass, n);
	b.w = mul_add(b.w, j, angJ);
}

template <class T>
MX_HD void velocity_impulse(Body<T>& b, V3<T> p, V3<T> r)
{
	typedef typename Num<T>::F F;
	if (val(b.invMass) == (F)0)
		return;
	b.v = mul_add(b.v, b.invMass, p);
	b.w = add(b.w, mat3_mulv(b.iw, cross(r, p)));
}

// SolveContactVelocity (src/xpbd.cpp:275-407): dynamic friction, restitution
// (applied unconditionally), rolling resistance. pv*/pw* are the bodies'
// velocities at the substep start, slip the points' slip directions.
template <class T>
MX_HD void velocity_op(const ContactGeo& g, Body<T>& a, Body<T>& b, V3<T> pvA, V3<T> pwA, V3<T> pvB, V3<T> pwB, T mu,
					   const T* lambdaNs, const PointPrep<T>* pp, float4* slip, float h, float gravityLen)
{
	typedef typename Num<T>::F F;
	V3<T> n = lift<T>(g.n);
	float e0 = g.mat.y, rr = g.mat.z;
	bool curved = (f2u(g.mat.w) & CF_CURVED) != 0;
	float invH = h > 0.0f ? 1.0f / h : 0.0f;
	T lambdaNTotal = Num<T>::c(0.0);
MX_UNROLL_POINTS
	for (int j = 0; j < 4; ++j)
	{
		if (j >= g.count)
			break;
		T lambdaN = lambdaNs[j];
		if (val(lambdaN) <= (F)0)
			continue;
		V3<T> rAf, rBf;
		if (curved)
		{
			rAf = contact_arm(g.tA, g.radiusA, g.hhA, a.q, g.pa(j), n, 1.0f);
			rBf = contact_arm(g.tB, g.radiusB, g.hhB, b.q, g.pb(j), n, -1.0f);
		}
		else
		{
			rAf = pp[j].rA;
			rBf = pp[j].rB;
		}
		V3<T> vrel = sub(add(b.v, cross(b.w, rBf)), add(a.v, cross(a.w, rAf)));
		T vn = dot(vrel, n);
		V3<T> vt = sub(vrel, scale(vn, n));
		T vtLen = length(vt);
		if (val(vtLen) > (F)1e-9)
		{
			V3<T> t = scale((F)1 / vtLen, vt);
			T wt = generalized_inv_mass(a, rAf, t) + generalized_inv_mass(b, rBf, t);
			if (val(wt) > (F)0)
			{
				const F staticSlip = (F)0.08;
				float4 sl = slip[j];
				bool coherent = val(t.x) * (F)sl.x + val(t.y) * (F)sl.y + val(t.z) * (F)sl.z > (F)0.7;
				T pStop = vtLen / wt;
				T pMax = mu * lambdaN * (F)invH;
				T P = !coherent && val(vtLen) < staticSlip ? pStop : (val(pStop) < val(pMax) ? pStop : pMax);
				slip[j] = make_float4((float)val(t.x), (float)val(t.y), (float)val(t.z), 0.0f);
				V3<T> imp = scale(-P, t);
				velocity_impulse(b, imp, rBf);
				velocity_impulse(a, neg(imp), rAf);
			}
		}
		V3<T> vrelPrev = sub(add(pvB, cross(pwB, rBf)), add(pvA, cross(pwA, rAf)));
		T vnPrev = dot(vrelPrev, n);
		F e = val(fabs_(vnPrev)) > (F)(2.0f * gravityLen * h) ? (F)e0 : (F)0;
		vrel = sub(add(b.v, cross(b.w, rBf)), add(a.v, cross(a.w, rAf)));
		vn = dot(vrel, n);
		T target = -e * vnPrev;
		target = val(target) > (F)0 ? target : Num<T>::c(0.0);
		T wN = pp[j].wN;
		if (val(wN) > (F)0)
		{
			T jj = (target - vn) / wN;
			axis_velocity(b, jj, n, pp[j].jB);
			axis_velocity(a, -jj, n, pp[j].jA);
		}
		lambdaNTotal = lambdaNTotal + (val(lambdaN) > (F)0 ? lambdaN : Num<T>::c(0.0));
	}
	if (rr > 0.0f && val(lambdaNTotal) > (F)0)
	{
		V3<T> wrel = sub(b.w, a.w);
		V3<T> wroll = sub(wrel, scale(dot(wrel, n), n));
		T wl = length(wroll);
		if (val(wl) > (F)1e-6)
		{
			V3<T> axis = scale((F)1 / wl, wroll);
			T k = dot(axis, mat3_mulv(a.iw, axis)) + dot(axis, mat3_mulv(b.iw, axis));
			if (val(k) > (F)1e-12)
			{
				T maxDw = (F)rr * lambdaNTotal * (F)invH * k;
				T dw = val(maxDw) < val(wl) ? maxDw : wl;
				V3<T> angImp = scale(dw / k, axis);
				b.w = sub(b.w, mat3_mulv(b.iw, angImp));
				a.w = add(a.w, mat3_mulv(a.iw, angImp));
			}
		}
	}
}

// ------------------------------------------------------------ world replay

template <class T>
MX_HD void integrate(const StepIn& in, StepState<T>& S, int32_t i, float h)
{
	if (!is_dynamic(in, i))
		return;
	S.px[i] = S.x[i];
	S.pq[i] = S.q[i];
	S.pv[i] = S.v[i];
	S.pw[i] = S.w[i];
	V3<T> x = S.x[i], v = S.v[i], w = S.w[i];
	Q4<T> q = S.q[i];
	M3<T> iw;
	integrate_op(x, q, v, w, S.il[i], iw, h, in.vel[i].w, in.invIL[i].w, in.angDamp[i], in.gravity);
	S.iw[i] = iw;
	S.x[i] = x;
	S.q[i] = q;
	S.v[i] = v;
	S.w[i] = w;
}

template <class T>
MX_HD Body<T> load_body(const StepState<T>& S, int32_t i)
{
	Body<T> b;
	b.p = S.x[i];
	b.q = S.q[i];
	b.v = S.v[i];
	b.w = S.w[i];
	b.invMass = S.invMass[i];
	b.iw = S.iw[i];
	return b;
}

template <class T>
MX_HD void store_body(StepState<T>& S, int32_t i, const Body<T>& b, const StepIn& in)
{
	if (!is_dynamic(in, i))
		return; // static and kinematic bodies keep their state
	S.x[i] = b.p;
	S.q[i] = b.q;
	S.v[i] = b.v;
	S.w[i] = b.w;
}

// freeze arms, Jacobians and normal mass; the manifold is the recorded one
// (no re-collision)
template <class T>
MX_HD void prepare(const StepIn& in, StepState<T>& S, int32_t c)
{
	int2 ab = in.bodies[c];
	ContactGeo g = contact_geo(in, c);
	for (int j = 0; j < g.count; ++j)
	{
		int32_t p = c * 4 + j;
		PointPrep<T> r = prepare_point(g, j, S.q[ab.x], S.q[ab.y], S.iw[ab.x], S.iw[ab.y], S.invMass[ab.x],
									   S.invMass[ab.y]);
		S.wN[p] = r.wN;
		S.jA[p] = r.jA;
		S.jB[p] = r.jB;
		S.rA[p] = r.rA;
		S.rB[p] = r.rB;
		S.lambdaN[p] = Num<T>::c(0.0);
		S.lambdaT[p] = Num<T>::c(0.0);
	}
}

template <class T>
MX_HD void normals(const StepIn& in, StepState<T>& S, int32_t c, float maxCorrection)
{
	int2 ab = in.bodies[c];
	int32_t ia = ab.x, ib = ab.y;
	ContactGeo g = contact_geo(in, c);
	V3<T> pa = S.x[ia], pb = S.x[ib];
	Q4<T> qa = S.q[ia], qb = S.q[ib];
	T lN[4], wN[4];
	V3<T> jA[4], jB[4];
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = c * 4 + (j < g.count ? j : 0);
		lN[j] = S.lambdaN[p];
		wN[j] = S.wN[p];
		jA[j] = S.jA[p];
		jB[j] = S.jB[p];
	}
	normals_any(in, c, g, pa, qa, pb, qb, S.invMass[ia], S.invMass[ib], lN, wN, jA, jB, maxCorrection);
	for (int j = 0; j < g.count; ++j)
		S.lambdaN[c * 4 + j] = lN[j];
	if (g.dynA)
	{
		S.x[ia] = pa;
		S.q[ia] = qa;
	}
	if (g.dynB)
	{
		S.x[ib] = pb;
		S.q[ib] = qb;
	}
}

template <class T>
MX_HD void friction(const StepIn& in, StepState<T>& S, int32_t c)
{
	int2 ab = in.bodies[c];
	ContactGeo g = contact_geo(in, c);
	if (g.count == 0)
		return;
	Body<T> a = load_body(S, ab.x), b = load_body(S, ab.y);
	T lN[4], lT[4];
	float4 fA[4], fB[4];
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = c * 4 + (j < g.count ? j : 0);
		lN[j] = S.lambdaN[p];
		lT[j] = S.lambdaT[p];
		fA[j] = S.fA[p];
		fB[j] = S.fB[p];
	}
	friction_op(g, a, b, S.mu[c], lN, lT, S.rA[c * 4], S.rB[c * 4], fA, fB);
	for (int j = 0; j < g.count; ++j)
	{
		S.lambdaT[c * 4 + j] = lT[j];
		S.fA[c * 4 + j] = fA[j];
		S.fB[c * 4 + j] = fB[j];
	}
	store_body(S, ab.x, a, in);
	store_body(S, ab.y, b, in);
}

template <class T>
MX_HD void reconstruct(const StepIn& in, StepState<T>& S, int32_t i, float invH)
{
	if (!is_dynamic(in, i))
		return;
	V3<T> v, w;
	reconstruct_op(S.x[i], S.q[i], S.px[i], S.pq[i], invH, v, w);
	S.v[i] = v;
	S.w[i] = w;
}

template <class T>
MX_HD void velocity(const StepIn& in, StepState<T>& S, int32_t c, float h, float gravityLen)
{
	int2 ab = in.bodies[c];
	int32_t ia = ab.x, ib = ab.y;
	ContactGeo g = contact_geo(in, c);
	Body<T> a = load_body(S, ia), b = load_body(S, ib);
	T lN[4];
	PointPrep<T> pp[4];
	float4 slip[4];
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = c * 4 + (j < g.count ? j : 0);
		lN[j] = S.lambdaN[p];
		pp[j].rA = S.rA[p];
		pp[j].rB = S.rB[p];
		pp[j].jA = S.jA[p];
		pp[j].jB = S.jB[p];
		pp[j].wN = S.wN[p];
		slip[j] = S.slip[p];
	}
	velocity_op(g, a, b, S.pv[ia], S.pw[ia], S.pv[ib], S.pw[ib], S.mu[c], lN, pp, slip, h, gravityLen);
	for (int j = 0; j < g.count; ++j)
		S.slip[c * 4 + j] = slip[j];
	store_body(S, ia, a, in);
	store_body(S, ib, b, in);
}

// One substep (src/xpbd.cpp:898-974 serial, :1269-1368 colored) with the
// manifold block of `in`: integrate, prepare, iterations x normals,
// stiction, reconstruct, velocity; the contact passes in the step's pass
// order (sched_keys).
template <class T>
MX_HD void solve_substep(const StepIn& in, StepState<T>& S, int iterations, float h, float invH, float maxCorrection,
						 float gravityLen)
{
	const int keys = sched_keys(in);
	for (int32_t c = 0; c < in.nc; ++c)
		for (int j = 0; j < 4; ++j)
		{
			S.fA[c * 4 + j] = in.pFA[c * 4 + j];
			S.fB[c * 4 + j] = in.pFB[c * 4 + j];
			S.slip[c * 4 + j] = in.pSlip[c * 4 + j];
		}
	for (int32_t i = 0; i < in.nb; ++i)
		integrate(in, S, i, h);
	for (int32_t c = 0; c < in.nc; ++c)
		prepare(in, S, c);
	for (int it = 0; it < iterations; ++it)
		for (int k = 0; k < keys; ++k)
			for (int32_t c = 0; c < in.nc; ++c)
				if (sched_key(in, c) == k)
					normals(in, S, c, maxCorrection);
	for (int k = 0; k < keys; ++k)
		for (int32_t c = 0; c < in.nc; ++c)
		{
			int2 ab = in.bodies[c];
			if (sched_key(in, c) == k && bf_shape(in.flags[ab.x]) == M3D_SHAPE_BOX &&
				bf_shape(in.flags[ab.y]) == M3D_SHAPE_BOX)
				friction(in, S, c);
		}
	for (int32_t i = 0; i < in.nb; ++i)
		reconstruct(in, S, i, invH);
	for (int k = 0; k < keys; ++k)
		for (int32_t c = 0; c < in.nc; ++c)
			if (sched_key(in, c) == k)
				velocity(in, S, c, h, gravityLen);
}

// The manifold block of substep s (blocks of cStride contacts).
MX_HD StepIn substep_in(const StepIn& in, int s)
{
	StepIn ins = in;
	ins.normal = in.normal + s * in.cStride;
	ins.cmat = in.cmat + s * in.cStride;
	ins.pA = in.pA + 4 * s * in.cStride;
	ins.pB = in.pB + 4 * s * in.cStride;
	ins.pFA = in.pFA + 4 * s * in.cStride;
	ins.pFB = in.pFB + 4 * s * in.cStride;
	ins.pSlip = in.pSlip + 4 * s * in.cStride;
	if (in.sched != nullptr)
		ins.sched = in.sched + s * in.cStride;
	return ins;
}

// The step's solve: its substeps, each with its recorded manifold block.
template <class T>
MX_HD void solve_step(const StepIn& in, StepState<T>& S, int iterations)
{
	float h = in.dt / (float)in.substeps;
	float invH = 1.0f / h;
	float maxCorrection = kMaxDepenetrationSpeed * h;
	float gravityLen = m3d_length(in.gravity);
	for (int s = 0; s < in.substeps; ++s)
		solve_substep(substep_in(in, s), S, iterations, h, invH, maxCorrection, gravityLen);
}

} // namespace dd
} // namespace mx
