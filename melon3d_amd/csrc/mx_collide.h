// Narrowphase (contact manifolds), friction-anchor inheritance, shape ray
// casts and shape AABBs, compiled for host and device.
//
// Each routine evaluates the same geometric algorithm, with the same
// floating-point expressions, as the reference routine cited beside it, so
// manifolds (points, feature ids, separations) match the reference bit for
// bit under strict IEEE arithmetic. Manifold normals point from A to B.
#pragma once

#include "mx_common.h"

namespace mx
{

// src/collide.cpp:16-34
MX_HD void add_point(Manifold* m, m3d_transform xfA, m3d_transform xfB, m3d_vec3 onA, m3d_vec3 onB, float sep,
					 uint32_t id)
{
	if (m->pointCount >= 4)
	{
		return;
	}
	MPoint mp;
	mp.localAnchorA = m3d_inv_transform_point(xfA, onA);
	mp.localAnchorB = m3d_inv_transform_point(xfB, onB);
	mp.frictionAnchorA = mp.localAnchorA;
	mp.frictionAnchorB = mp.localAnchorB;
	mp.separation = sep;
	mp.id = id;
	mp.lambdaN = 0.0f;
	mp.lambdaT = 0.0f;
	mp.separationOffset = 0.0f;
	mp.slipDir = m3d_v3_zero();
	m->set_point(m->pointCount++, mp);
}

// src/collide.cpp:36-47
MX_HD m3d_vec3 closest_on_segment(m3d_vec3 p, m3d_vec3 a, m3d_vec3 b)
{
	m3d_vec3 ab = m3d_sub(b, a);
	float dd = m3d_dot(ab, ab);
	if (dd < 1e-12f)
	{
		return a;
	}
	float t = m3d_dot(m3d_sub(p, a), ab) / dd;
	t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
	return m3d_mul_add(a, t, ab);
}

// Closest points of two segments (Ericson, RTCD 5.1.9); src/collide.cpp:50-101
MX_HD void closest_segment_segment(m3d_vec3 p1, m3d_vec3 q1, m3d_vec3 p2, m3d_vec3 q2, m3d_vec3* c1, m3d_vec3* c2)
{
	m3d_vec3 d1 = m3d_sub(q1, p1);
	m3d_vec3 d2 = m3d_sub(q2, p2);
	m3d_vec3 r = m3d_sub(p1, p2);
	float a = m3d_dot(d1, d1);
	float e = m3d_dot(d2, d2);
	float f = m3d_dot(d2, r);
	float s, t;
	if (a < 1e-12f && e < 1e-12f)
	{
		*c1 = p1;
		*c2 = p2;
		return;
	}
	if (a < 1e-12f)
	{
		s = 0.0f;
		t = f / e;
	}
	else
	{
		float c = m3d_dot(d1, r);
		if (e < 1e-12f)
		{
			t = 0.0f;
			s = -c / a;
		}
		else
		{
			float b = m3d_dot(d1, d2);
			float den = a * e - b * b;
			s = den > 1e-12f ? (b * f - c * e) / den : 0.0f;
			s = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
			t = (b * s + f) / e;
			if (t < 0.0f)
			{
				t = 0.0f;
				s = -c / a;
			}
			else if (t > 1.0f)
			{
				t = 1.0f;
				s = (b - c) / a;
			}
			s = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
		}
	}
	*c1 = m3d_mul_add(p1, s, d1);
	*c2 = m3d_mul_add(p2, t, d2);
}

// src/collide.cpp:103-110
MX_HD m3d_vec3 clamp_to_box(m3d_vec3 p, m3d_vec3 e)
{
	m3d_vec3 q;
	q.x = p.x < -e.x ? -e.x : (p.x > e.x ? e.x : p.x);
	q.y = p.y < -e.y ? -e.y : (p.y > e.y ? e.y : p.y);
	q.z = p.z < -e.z ? -e.z : (p.z > e.z ? e.z : p.z);
	return q;
}

MX_HD m3d_vec3 dir_or_up(m3d_vec3 d, float len)
{
	return len > 1e-9f ? m3d_scale(1.0f / len, d) : m3d_v3(0.0f, 1.0f, 0.0f);
}

// sphere A - sphere B; src/collide.cpp:116-130
MX_HD void collide_spheres(const Shape& a, m3d_transform xfA, const Shape& b, m3d_transform xfB, float margin,
						   Manifold* m)
{
	m3d_vec3 d = m3d_sub(xfB.p, xfA.p);
	float dist = m3d_length(d);
	m3d_vec3 n = dir_or_up(d, dist);
	float sep = dist - a.radius - b.radius;
	if (sep > margin)
	{
		return;
	}
	m->normal = n;
	add_point(m, xfA, xfB, m3d_mul_add(xfA.p, a.radius, n), m3d_mul_add(xfB.p, -b.radius, n), sep, 0);
}

// sphere A - capsule B; src/collide.cpp:136-155
MX_HD void collide_sphere_capsule(const Shape& a, m3d_transform xfA, const Shape& b, m3d_transform xfB, float margin,
								  Manifold* m)
{
	m3d_vec3 axis = m3d_rotate(xfB.q, m3d_v3(0.0f, b.halfHeight, 0.0f));
	m3d_vec3 c = closest_on_segment(xfA.p, m3d_sub(xfB.p, axis), m3d_add(xfB.p, axis));
	m3d_vec3 d = m3d_sub(c, xfA.p);
	float dist = m3d_length(d);
	m3d_vec3 n = dir_or_up(d, dist);
	float sep = dist - a.radius - b.radius;
	if (sep > margin)
	{
		return;
	}
	m->normal = n;
	add_point(m, xfA, xfB, m3d_mul_add(xfA.p, a.radius, n), m3d_mul_add(c, -b.radius, n), sep, 0);
}

// capsule A - capsule B; src/collide.cpp:161-230
MX_HD void collide_capsules(const Shape& a, m3d_transform xfA, const Shape& b, m3d_transform xfB, float margin,
							Manifold* m)
{
	m3d_vec3 axisA = m3d_rotate(xfA.q, m3d_v3(0.0f, a.halfHeight, 0.0f));
	m3d_vec3 a1 = m3d_sub(xfA.p, axisA);
	m3d_vec3 a2 = m3d_add(xfA.p, axisA);
	m3d_vec3 axisB = m3d_rotate(xfB.q, m3d_v3(0.0f, b.halfHeight, 0.0f));
	m3d_vec3 b1 = m3d_sub(xfB.p, axisB);
	m3d_vec3 b2 = m3d_add(xfB.p, axisB);
	m3d_vec3 uA = m3d_normalize(m3d_sub(a2, a1));
	m3d_vec3 uB = m3d_normalize(m3d_sub(b2, b1));
	float par = m3d_length(m3d_cross(uA, uB));

	// near-parallel cores: two points over the shared interval
	if (par < 0.05f && a.halfHeight > 1e-6f)
	{
		float lenA = 2.0f * a.halfHeight;
		float t1 = m3d_dot(m3d_sub(b1, a1), uA);
		float t2 = m3d_dot(m3d_sub(b2, a1), uA);
		float lo = fmaxf(0.0f, fminf(t1, t2));
		float hi = fminf(lenA, fmaxf(t1, t2));
		if (hi - lo > 1e-4f)
		{
			m3d_vec3 midA = m3d_mul_add(a1, 0.5f * (lo + hi), uA);
			m3d_vec3 midB = closest_on_segment(midA, b1, b2);
			m3d_vec3 d = m3d_sub(midB, midA);
			float dist = m3d_length(d);
			m3d_vec3 n = dir_or_up(d, dist);
			if (dist - a.radius - b.radius > margin)
			{
				return;
			}
			m->normal = n;
			for (int i = 0; i < 2; ++i)
			{
				float ti = i == 0 ? lo : hi;
				m3d_vec3 cA = m3d_mul_add(a1, ti, uA);
				m3d_vec3 cB = closest_on_segment(cA, b1, b2);
				float di = m3d_length(m3d_sub(cB, cA));
				float sep = di - a.radius - b.radius;
				if (sep <= margin)
				{
					add_point(m, xfA, xfB, m3d_mul_add(cA, a.radius, n), m3d_mul_add(cB, -b.radius, n), sep,
							  (uint32_t)i);
				}
			}
			if (m->pointCount > 0)
			{
				return;
			}
		}
	}

	m3d_vec3 cA, cB;
	closest_segment_segment(a1, a2, b1, b2, &cA, &cB);
	m3d_vec3 d = m3d_sub(cB, cA);
	float dist = m3d_length(d);
	m3d_vec3 n = dir_or_up(d, dist);
	float sep = dist - a.radius - b.radius;
	if (sep > margin)
	{
		return;
	}
	m->normal = n;
	add_point(m, xfA, xfB, m3d_mul_add(cA, a.radius, n), m3d_mul_add(cB, -b.radius, n), sep, 0);
}

// Least-penetration face of a box for a point c inside it (sphere-box and
// capsule-box share the tie-breaking order: x, then y, then z).
MX_HD int least_face(m3d_vec3 e, m3d_vec3 c, m3d_vec3* nLocal, float* depth)
{
	float dx = e.x - fabsf(c.x);
	float dy = e.y - fabsf(c.y);
	float dz = e.z - fabsf(c.z);
	if (dx < dy && dx < dz)
	{
		*nLocal = m3d_v3(c.x >= 0.0f ? 1.0f : -1.0f, 0.0f, 0.0f);
		*depth = dx;
		return 0;
	}
	if (dy < dz)
	{
		*nLocal = m3d_v3(0.0f, c.y >= 0.0f ? 1.0f : -1.0f, 0.0f);
		*depth = dy;
		return 1;
	}
	*nLocal = m3d_v3(0.0f, 0.0f, c.z >= 0.0f ? 1.0f : -1.0f);
	*depth = dz;
	return 2;
}

// sphere A - box B; src/collide.cpp:236-288
MX_HD void collide_sphere_box(const Shape& a, m3d_transform xfA, const Shape& b, m3d_transform xfB, float margin,
							  Manifold* m)
{
	m3d_vec3 c = m3d_inv_transform_point(xfB, xfA.p);
	m3d_vec3 e = b.halfExtents;
	m3d_vec3 q = clamp_to_box(c, e);
	m3d_vec3 d = m3d_sub(c, q);
	float d2 = m3d_length_sq(d);
	m3d_vec3 nLocal;
	float sep;
	if (d2 > 1e-12f)
	{
		float dist = sqrtf(d2);
		nLocal = m3d_scale(1.0f / dist, d);
		sep = dist - a.radius;
	}
	else
	{
		float depth;
		int axis = least_face(e, c, &nLocal, &depth);
		sep = -(depth + a.radius);
		q = c;
		v3_set(q, axis, v3_get(nLocal, axis) * v3_get(e, axis));
	}
	if (sep > margin)
	{
		return;
	}
	m->normal = m3d_neg(m3d_rotate(xfB.q, nLocal));
	m3d_vec3 pB = m3d_transform_point(xfB, q);
	m3d_vec3 pA = m3d_mul_add(xfA.p, a.radius, m->normal);
	add_point(m, xfA, xfB, pA, pB, sep, 0);
}

// src/collide.cpp:295-299
MX_HD float point_box_dist_sq(m3d_vec3 p, m3d_vec3 e)
{
	return m3d_length_sq(m3d_sub(p, clamp_to_box(p, e)));
}

// capsule A - box B; src/collide.cpp:301-424
MX_HD void collide_capsule_box(const Shape& a, m3d_transform xfA, const Shape& b, m3d_transform xfB, float margin,
							   Manifold* m)
{
	m3d_vec3 e = b.halfExtents;
	float r = a.radius;
	m3d_vec3 axis = m3d_rotate(xfA.q, m3d_v3(0.0f, a.halfHeight, 0.0f));
	m3d_vec3 s0 = m3d_inv_transform_point(xfB, m3d_sub(xfA.p, axis));
	m3d_vec3 s1 = m3d_inv_transform_point(xfB, m3d_add(xfA.p, axis));
	m3d_vec3 dir = m3d_sub(s1, s0);

	// distance to a convex set is convex along the segment: ternary search
	float lo = 0.0f, hi = 1.0f;
	float len2 = m3d_length_sq(dir);
	float tol = len2 > 1e-12f ? (kLinearSlop * kLinearSlop) / len2 : 1e-6f;
	for (int it = 0; it < 22; ++it)
	{
		if ((hi - lo) * (hi - lo) < tol)
		{
			break;
		}
		float t1 = lo + (hi - lo) / 3.0f;
		float t2 = hi - (hi - lo) / 3.0f;
		float d1 = point_box_dist_sq(m3d_mul_add(s0, t1, dir), e);
		float d2 = point_box_dist_sq(m3d_mul_add(s0, t2, dir), e);
		if (d1 < d2)
		{
			hi = t2;
		}
		else
		{
			lo = t1;
		}
	}
	float ts = 0.5f * (lo + hi);
	m3d_vec3 c = m3d_mul_add(s0, ts, dir);
	m3d_vec3 q = clamp_to_box(c, e);
	m3d_vec3 diff = m3d_sub(c, q);
	float dist2 = m3d_length_sq(diff);
	m3d_vec3 nLocal;
	float sep;
	if (dist2 > 1e-10f)
	{
		float dist = sqrtf(dist2);
		nLocal = m3d_scale(1.0f / dist, diff);
		sep = dist - r;
		if (sep > margin)
		{
			return;
		}
	}
	else
	{
		float depth;
		least_face(e, c, &nLocal, &depth);
		sep = -(depth + r);
		q = m3d_sub(c, m3d_scale(sep + r, nLocal));
	}
	m->normal = m3d_neg(m3d_rotate(xfB.q, nLocal));

	// capsule lying flat on the face: one point per segment end
	m3d_vec3 u = m3d_normalize(dir);
	if (fabsf(m3d_dot(u, nLocal)) < 0.05f && a.halfHeight > 1e-6f)
	{
		int added = 0;
		for (int i = 0; i < 2; ++i)
		{
			m3d_vec3 p = i == 0 ? s0 : s1;
			m3d_vec3 qp = clamp_to_box(p, e);
			m3d_vec3 dp = m3d_sub(p, qp);
			float d2 = m3d_length_sq(dp);
			float sp;
			if (d2 > 1e-10f)
			{
				sp = sqrtf(d2) - r;
			}
			else
			{
				float ext = e.x * fabsf(nLocal.x) + e.y * fabsf(nLocal.y) + e.z * fabsf(nLocal.z);
				sp = (m3d_dot(p, nLocal) - ext) - r;
				qp = m3d_sub(p, m3d_scale(m3d_dot(p, nLocal) - ext, nLocal));
			}
			if (sp <= margin)
			{
				m3d_vec3 onCap = m3d_mul_add(m3d_transform_point(xfB, p), r, m->normal);
				add_point(m, xfA, xfB, onCap, m3d_transform_point(xfB, qp), sp, (uint32_t)i);
				++added;
			}
		}
		if (added > 0)
		{
			return;
		}
	}
	m3d_vec3 onCap = m3d_mul_add(m3d_transform_point(xfB, c), r, m->normal);
	add_point(m, xfA, xfB, onCap, m3d_transform_point(xfB, q), sep, 4);
}

// ------------------------------------------------------------ box - box

struct ClipV
{
	m3d_vec3 v;
	uint32_t id;
};

// A clipping polygon of up to 8 vertices kept in vector registers (an
// element taken at a run-time index is a register lane, not a memory slot).
typedef float mx_float8 __attribute__((ext_vector_type(8)));
typedef uint32_t mx_uint8 __attribute__((ext_vector_type(8)));
struct ClipPoly
{
	mx_float8 x, y, z;
	mx_uint8 id;
	MX_HD ClipV get(int i) const
	{
		ClipV c;
		c.v = m3d_v3(x[i], y[i], z[i]);
		c.id = id[i];
		return c;
	}
	MX_HD void put(int i, const ClipV& c)
	{
		x[i] = c.v.x;
		y[i] = c.v.y;
		z[i] = c.v.z;
		id[i] = c.id;
	}
};

// Column i of a rotation (a box's axis i) at a run-time index.
MX_HD m3d_vec3 mat3_col(const m3d_mat3& m, int i) { return i == 0 ? m.cx : (i == 1 ? m.cy : m.cz); }

// Sutherland-Hodgman against dot(x, n) <= off, at most 8 output vertices;
// src/collide.cpp:437-464
MX_HD int clip_polygon(const ClipPoly& in, int inCount, m3d_vec3 n, float off, uint32_t planeId, ClipPoly& out)
{
	int oc = 0;
	for (int i = 0; i < inCount; ++i)
	{
		const ClipV v1 = in.get(i);
		const ClipV v2 = in.get(i + 1 < inCount ? i + 1 : 0);
		float d1 = m3d_dot(v1.v, n) - off;
		float d2 = m3d_dot(v2.v, n) - off;
		if (d1 <= 0.0f)
		{
			out.put(oc++, v1);
		}
		if (d1 * d2 < 0.0f && oc < 8)
		{
			float t = d1 / (d1 - d2);
			ClipV x;
			x.v = m3d_lerp(v1.v, v2.v, t);
			x.id = ((v1.id ^ v2.id) << 8) | planeId;
			out.put(oc++, x);
		}
		if (oc >= 8)
		{
			break;
		}
	}
	return oc;
}

// Support edge of a box along a world direction; src/collide.cpp:467-509
MX_HD void support_edge(m3d_transform xf, const m3d_mat3& rot, m3d_vec3 e, m3d_vec3 dirWorld, m3d_vec3* pOut,
						m3d_vec3* qOut)
{
	m3d_vec3 d = m3d_mat3_mulv_t(&rot, dirWorld);
	m3d_vec3 ad = m3d_abs(d);
	int ax = 0;
	if (ad.y < ad.x && ad.y <= ad.z)
	{
		ax = 1;
	}
	else if (ad.z < ad.x && ad.z <= ad.y)
	{
		ax = 2;
	}
	m3d_vec3 p = m3d_v3(d.x >= 0.0f ? e.x : -e.x, d.y >= 0.0f ? e.y : -e.y, d.z >= 0.0f ? e.z : -e.z);
	m3d_vec3 q = p;
	v3_set(p, ax, -v3_get(e, ax));
	v3_set(q, ax, v3_get(e, ax));
	*pOut = m3d_transform_point(xf, p);
	*qOut = m3d_transform_point(xf, q);
}

MX_HD float project_radius(const m3d_vec3* axes, m3d_vec3 ext, m3d_vec3 n)
{
	return ext.x * fabsf(m3d_dot(axes[0], n)) + ext.y * fabsf(m3d_dot(axes[1], n)) + ext.z * fabsf(m3d_dot(axes[2], n));
}

// SAT over 15 axes, then edge point or reference-face clipping;
// src/collide.cpp:511-805
MX_HD void collide_boxes(const Shape& a, m3d_transform xfA, const Shape& b, m3d_transform xfB, float margin,
						 Manifold* m)
{
	m3d_mat3 rotA = m3d_quat_to_mat3(xfA.q);
	m3d_mat3 rotB = m3d_quat_to_mat3(xfB.q);
	m3d_vec3 eA = a.halfExtents;
	m3d_vec3 eB = b.halfExtents;
	m3d_vec3 d = m3d_sub(xfB.p, xfA.p);
	const m3d_vec3 axA[3] = { rotA.cx, rotA.cy, rotA.cz };
	const m3d_vec3 axB[3] = { rotB.cx, rotB.cy, rotB.cz };

	float faceASep = -FLT_MAX;
	int faceA = 0;
	for (int i = 0; i < 3; ++i)
	{
		float s = fabsf(m3d_dot(d, axA[i])) - (v3_get(eA, i) + project_radius(axB, eB, axA[i]));
		if (s > margin)
		{
			return;
		}
		if (s > faceASep)
		{
			faceASep = s;
			faceA = i;
		}
	}
	float faceBSep = -FLT_MAX;
	int faceB = 0;
	for (int i = 0; i < 3; ++i)
	{
		float s = fabsf(m3d_dot(d, axB[i])) - (v3_get(eB, i) + project_radius(axA, eA, axB[i]));
		if (s > margin)
		{
			return;
		}
		if (s > faceBSep)
		{
			faceBSep = s;
			faceB = i;
		}
	}
	float edgeSep = -FLT_MAX;
	int edgeI = 0, edgeJ = 0;
	m3d_vec3 edgeAxis = m3d_v3_zero();
	for (int i = 0; i < 3; ++i)
	{
		for (int j = 0; j < 3; ++j)
		{
			m3d_vec3 ax = m3d_cross(axA[i], axB[j]);
			float l2 = m3d_length_sq(ax);
			if (l2 < 1e-8f)
			{
				continue;
			}
			ax = m3d_scale(1.0f / sqrtf(l2), ax);
			float s = fabsf(m3d_dot(d, ax)) - (project_radius(axA, eA, ax) + project_radius(axB, eB, ax));
			if (s > margin)
			{
				return;
			}
			if (s > edgeSep)
			{
				edgeSep = s;
				edgeI = i;
				edgeJ = j;
				edgeAxis = ax;
			}
		}
	}

	// face preference tolerances; src/collide.cpp:589-602
	const float relTol = 0.95f;
	const float absTol = 0.01f * kLinearSlop * 200.0f;
	float bestFace = faceASep;
	bool refIsA = true;
	int refFace = faceA;
	if (faceBSep > relTol * faceASep + absTol)
	{
		bestFace = faceBSep;
		refIsA = false;
		refFace = faceB;
	}

	if (edgeSep > relTol * bestFace + absTol)
	{
		m3d_vec3 n = edgeAxis;
		if (m3d_dot(d, n) < 0.0f)
		{
			n = m3d_neg(n);
		}
		m3d_vec3 pA1, pA2, pB1, pB2;
		support_edge(xfA, rotA, eA, n, &pA1, &pA2);
		support_edge(xfB, rotB, eB, m3d_neg(n), &pB1, &pB2);
		m3d_vec3 cA, cB;
		closest_segment_segment(pA1, pA2, pB1, pB2, &cA, &cB);
		m->normal = n;
		add_point(m, xfA, xfB, cA, cB, edgeSep, 0xE0000000u | (uint32_t)(edgeI * 3 + edgeJ));
		return;
	}

	// face contact: clip the incident face against the reference side planes
	// (the axes of either box are picked with compile-time indices)
	m3d_transform xfRef = refIsA ? xfA : xfB;
	m3d_transform xfInc = refIsA ? xfB : xfA;
	// picked component by component (a choice between two whole matrices
	// or vectors would put both in scratch memory)
	auto sel = [](bool c, m3d_vec3 a, m3d_vec3 b) { return m3d_v3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); };
	m3d_mat3 rotInc, rotRef;
	rotInc.cx = sel(refIsA, rotB.cx, rotA.cx);
	rotInc.cy = sel(refIsA, rotB.cy, rotA.cy);
	rotInc.cz = sel(refIsA, rotB.cz, rotA.cz);
	rotRef.cx = sel(refIsA, rotA.cx, rotB.cx);
	rotRef.cy = sel(refIsA, rotA.cy, rotB.cy);
	rotRef.cz = sel(refIsA, rotA.cz, rotB.cz);
	auto refAx = [&](int i) { return mat3_col(rotRef, i); };
	auto incAx = [&](int i) { return mat3_col(rotInc, i); };
	m3d_vec3 refExt = refIsA ? eA : eB;
	m3d_vec3 incExt = refIsA ? eB : eA;
	m3d_vec3 dRI = refIsA ? d : m3d_neg(d);

	const m3d_vec3 refAxis = refAx(refFace);
	float refSign = m3d_dot(refAxis, dRI) >= 0.0f ? 1.0f : -1.0f;
	m3d_vec3 refN = m3d_scale(refSign, refAxis);

	int incFace = 0;
	float minDot = FLT_MAX;
	float incSign = 1.0f;
#pragma unroll
	for (int i = 0; i < 3; ++i)
	{
		float dt = m3d_dot(incAx(i), refN);
		if (dt < minDot)
		{
			minDot = dt;
			incFace = i;
			incSign = 1.0f;
		}
		if (-dt < minDot)
		{
			minDot = -dt;
			incFace = i;
			incSign = -1.0f;
		}
	}

	int iu = incFace + 1 < 3 ? incFace + 1 : incFace - 2;
	int iv = incFace + 2 < 3 ? incFace + 2 : incFace - 1;
	m3d_vec3 cL = m3d_v3_zero(), uL = m3d_v3_zero(), vL = m3d_v3_zero();
	v3_set(cL, incFace, incSign * v3_get(incExt, incFace));
	v3_set(uL, iu, v3_get(incExt, iu));
	v3_set(vL, iv, v3_get(incExt, iv));

	ClipPoly poly, tmp;
	{
		m3d_vec3 cW = m3d_transform_point(xfInc, cL);
		m3d_vec3 uW = m3d_mat3_mulv(&rotInc, uL);
		m3d_vec3 vW = m3d_mat3_mulv(&rotInc, vL);
		ClipV q;
		q.v = m3d_add(m3d_add(cW, uW), vW);
		q.id = 0;
		poly.put(0, q);
		q.v = m3d_add(m3d_sub(cW, uW), vW);
		q.id = 1;
		poly.put(1, q);
		q.v = m3d_sub(m3d_sub(cW, uW), vW);
		q.id = 2;
		poly.put(2, q);
		q.v = m3d_sub(m3d_add(cW, uW), vW);
		q.id = 3;
		poly.put(3, q);
	}
	int count = 4;
	int ru = refFace + 1 < 3 ? refFace + 1 : refFace - 2;
	int rv = refFace + 2 < 3 ? refFace + 2 : refFace - 1;
#pragma unroll
	for (int k = 0; k < 2; ++k)
	{
		int ai = k == 0 ? ru : rv;
		m3d_vec3 sn = refAx(ai);
		float cp = m3d_dot(xfRef.p, sn);
		count = clip_polygon(poly, count, sn, cp + v3_get(refExt, ai), (uint32_t)(2 * k + 4), tmp);
		if (count == 0)
		{
			return;
		}
		count = clip_polygon(tmp, count, m3d_neg(sn), -(cp - v3_get(refExt, ai)), (uint32_t)(2 * k + 5), poly);
		if (count == 0)
		{
			return;
		}
	}

	float faceOff = m3d_dot(xfRef.p, refN) + v3_get(refExt, refFace);
	ClipPoly kept;
	mx_float8 seps;
	int kc = 0;
	for (int i = 0; i < count; ++i)
	{
		const ClipV pv = poly.get(i);
		float s = m3d_dot(pv.v, refN) - faceOff;
		if (s <= margin)
		{
			kept.put(kc, pv);
			seps[kc] = s;
			++kc;
		}
	}
	if (kc == 0)
	{
		return;
	}

	m->normal = refIsA ? refN : m3d_neg(refN);
	const uint32_t flip = refIsA ? 0u : 0x40000000u;
	// one kept vertex as a manifold point: the incident vertex and its
	// projection on the reference face
	auto emit = [&](const ClipV& kv, float sep) {
		m3d_vec3 pInc = kv.v;
		m3d_vec3 pRef = m3d_mul_add(pInc, -sep, refN);
		uint32_t id = (kv.id & 0x00FFFFFFu) | ((uint32_t)refFace << 24) | flip;
		if (refIsA)
		{
			add_point(m, xfA, xfB, pRef, pInc, sep, id);
		}
		else
		{
			add_point(m, xfA, xfB, pInc, pRef, sep, id);
		}
	};
	if (kc <= 4)
	{
		for (int i = 0; i < kc; ++i)
		{
			emit(kept.get(i), seps[i]);
		}
		return;
	}
	// reduce to 4: deepest, farthest from it, then extreme signed areas
	int i0 = 0;
	float s0 = seps[0];
	for (int i = 1; i < kc; ++i)
	{
		if (seps[i] < s0)
		{
			s0 = seps[i];
			i0 = i;
		}
	}
	const m3d_vec3 k0 = kept.get(i0).v;
	int i1 = -1;
	float best = -1.0f;
	for (int i = 0; i < kc; ++i)
	{
		if (i == i0)
		{
			continue;
		}
		float dd = m3d_length_sq(m3d_sub(kept.get(i).v, k0));
		if (dd > best)
		{
			best = dd;
			i1 = i;
		}
	}
	const m3d_vec3 e1 = m3d_sub(kept.get(i1).v, k0);
	int i2 = -1, i3 = -1;
	float bestArea = 0.0f, worstArea = 0.0f;
	for (int i = 0; i < kc; ++i)
	{
		if (i == i0 || i == i1)
		{
			continue;
		}
		m3d_vec3 e2 = m3d_sub(kept.get(i).v, k0);
		float area = m3d_dot(m3d_cross(e1, e2), refN);
		if (area > bestArea)
		{
			bestArea = area;
			i2 = i;
		}
		if (area < worstArea)
		{
			worstArea = area;
			i3 = i;
		}
	}
	emit(kept.get(i0), seps[i0]);
	emit(kept.get(i1), seps[i1]);
	if (i2 >= 0)
	{
		emit(kept.get(i2), seps[i2]);
	}
	if (i3 >= 0)
	{
		emit(kept.get(i3), seps[i3]);
	}
}

// Dispatch in canonical type order, flipping when A's type is higher;
// src/collide.cpp:813-859. Slots at or beyond the new point count keep their
// previous contents.
MX_HD void collide_shapes(const Shape& sA, m3d_transform xfA, const Shape& sB, m3d_transform xfB, float margin,
						  Manifold* m)
{
	m->pointCount = 0;
	m->normal = m3d_v3(0.0f, 1.0f, 0.0f);
	// canonical order (type A <= type B), then flipped back in place: the
	// same points as the reference's flipped copy (src/collide.cpp:818-836),
	// and slots beyond the count keep their old values either way
	bool flip = sA.type > sB.type;
	const Shape& s0 = flip ? sB : sA;
	const Shape& s1 = flip ? sA : sB;
	m3d_transform x0 = flip ? xfB : xfA;
	m3d_transform x1 = flip ? xfA : xfB;
	switch (s0.type)
	{
		case M3D_SHAPE_SPHERE:
			if (s1.type == M3D_SHAPE_SPHERE)
				collide_spheres(s0, x0, s1, x1, margin, m);
			else if (s1.type == M3D_SHAPE_CAPSULE)
				collide_sphere_capsule(s0, x0, s1, x1, margin, m);
			else
				collide_sphere_box(s0, x0, s1, x1, margin, m);
			break;
		case M3D_SHAPE_CAPSULE:
			if (s1.type == M3D_SHAPE_CAPSULE)
				collide_capsules(s0, x0, s1, x1, margin, m);
			else if (s1.type == M3D_SHAPE_BOX)
				collide_capsule_box(s0, x0, s1, x1, margin, m);
			break;
		default:
			collide_boxes(s0, x0, s1, x1, margin, m);
			break;
	}
	if (!flip)
	{
		return;
	}
	m->normal = m3d_neg(m->normal);
	for (int i = 0; i < m->pointCount; ++i)
	{
		MPoint p = m->point(i);
		m3d_vec3 t = p.localAnchorA;
		p.localAnchorA = p.localAnchorB;
		p.localAnchorB = t;
		t = p.frictionAnchorA;
		p.frictionAnchorA = p.frictionAnchorB;
		p.frictionAnchorB = t;
		m->set_point(i, p);
	}
}

// Stiction anchors survive regeneration for box-box pairs while the stored
// pair stretched < 2 cm and the feature drifted < 5 cm; the slip direction
// carries over for every shape. src/collide.cpp:865-913
template <class OldManifold>
MX_HD void inherit_friction_anchors(const OldManifold& old, Manifold& m, m3d_vec3 posA, m3d_quat rotA, m3d_vec3 posB,
									m3d_quat rotB, bool boxBox)
{
	const float slipRelease2 = 0.02f * 0.02f;
	const float featureDrift2 = 0.05f * 0.05f;
	m3d_vec3 n = m.normal;
	for (int i = 0; i < m.pointCount; ++i)
	{
		MPoint mp = m.point(i);
		for (int j = 0; j < old.pointCount; ++j)
		{
			const MPoint op = old.point(j);
			if (op.id != mp.id)
			{
				continue;
			}
			mp.slipDir = op.slipDir;
			if (!boxBox)
			{
				break;
			}
			m3d_vec3 pAold = m3d_add(posA, m3d_rotate(rotA, op.frictionAnchorA));
			m3d_vec3 pBold = m3d_add(posB, m3d_rotate(rotB, op.frictionAnchorB));
			m3d_vec3 pAnew = m3d_add(posA, m3d_rotate(rotA, mp.localAnchorA));
			m3d_vec3 gap = m3d_sub(pBold, pAold);
			m3d_vec3 dpt = m3d_sub(gap, m3d_scale(m3d_dot(gap, n), n));
			if (m3d_length_sq(dpt) < slipRelease2 && m3d_length_sq(m3d_sub(pAold, pAnew)) < featureDrift2)
			{
				mp.frictionAnchorA = op.frictionAnchorA;
				mp.frictionAnchorB = op.frictionAnchorB;
			}
			break;
		}
		m.set_point(i, mp);
	}
}

// ------------------------------------------------------------ ray casts

// src/collide.cpp:919-938
MX_HD bool ray_sphere(m3d_vec3 center, float radius, m3d_vec3 o, m3d_vec3 d, float* fraction, m3d_vec3* normal)
{
	m3d_vec3 s = m3d_sub(o, center);
	float b = m3d_length_sq(s) - radius * radius;
	float rr = m3d_dot(s, d);
	float c = m3d_dot(d, d);
	float sigma = rr * rr - c * b;
	if (sigma < 0.0f || c < 1e-12f)
	{
		return false;
	}
	float t = -(rr + sqrtf(sigma)) / c;
	if (t < 0.0f || t > 1.0f)
	{
		return false;
	}
	*fraction = t;
	*normal = m3d_normalize(m3d_add(s, m3d_scale(t, d)));
	return true;
}

// slab test in box space; src/collide.cpp:940-989
MX_HD bool ray_box_local(m3d_vec3 e, m3d_vec3 o, m3d_vec3 d, float* fraction, m3d_vec3* nLocal)
{
	float tmin = 0.0f, tmax = 1.0f;
	m3d_vec3 n = m3d_v3_zero();
	for (int i = 0; i < 3; ++i)
	{
		float oi = v3_get(o, i), di = v3_get(d, i), ei = v3_get(e, i);
		if (fabsf(di) < 1e-12f)
		{
			if (oi < -ei || oi > ei)
			{
				return false;
			}
		}
		else
		{
			float inv = 1.0f / di;
			float t1 = (-ei - oi) * inv;
			float t2 = (ei - oi) * inv;
			float sign = -1.0f;
			if (t1 > t2)
			{
				float tt = t1;
				t1 = t2;
				t2 = tt;
				sign = 1.0f;
			}
			if (t1 > tmin)
			{
				tmin = t1;
				n = m3d_v3_zero();
				v3_set(n, i, sign);
			}
			tmax = tmax < t2 ? tmax : t2;
			if (tmin > tmax)
			{
				return false;
			}
		}
	}
	if (m3d_length_sq(n) < 0.5f)
	{
		return false;
	}
	*fraction = tmin;
	*nLocal = n;
	return true;
}

// src/collide.cpp:991-1066
MX_HD bool ray_cast_shape(const Shape& s, m3d_transform xf, m3d_vec3 origin, m3d_vec3 tr, float* fraction,
						  m3d_vec3* normal)
{
	if (s.type == M3D_SHAPE_SPHERE)
	{
		return ray_sphere(xf.p, s.radius, origin, tr, fraction, normal);
	}
	m3d_vec3 oL = m3d_inv_transform_point(xf, origin);
	m3d_vec3 dL = m3d_inv_rotate(xf.q, tr);
	if (s.type == M3D_SHAPE_BOX)
	{
		m3d_vec3 nL;
		if (!ray_box_local(s.halfExtents, oL, dL, fraction, &nL))
		{
			return false;
		}
		*normal = m3d_rotate(xf.q, nL);
		return true;
	}
	float h = s.halfHeight;
	float r = s.radius;
	float best = FLT_MAX;
	m3d_vec3 bestN = m3d_v3_zero();
	float aa = dL.x * dL.x + dL.z * dL.z;
	if (aa > 1e-12f)
	{
		float bb = oL.x * dL.x + oL.z * dL.z;
		float cc = oL.x * oL.x + oL.z * oL.z - r * r;
		float sigma = bb * bb - aa * cc;
		if (sigma >= 0.0f)
		{
			float t = (-bb - sqrtf(sigma)) / aa;
			if (t >= 0.0f && t <= 1.0f)
			{
				float y = oL.y + t * dL.y;
				if (fabsf(y) <= h)
				{
					best = t;
					m3d_vec3 hp = m3d_mul_add(oL, t, dL);
					bestN = m3d_normalize(m3d_v3(hp.x, 0.0f, hp.z));
				}
			}
		}
	}
	for (int i = 0; i < 2; ++i)
	{
		float t;
		m3d_vec3 n;
		if (ray_sphere(m3d_v3(0.0f, i == 0 ? h : -h, 0.0f), r, oL, dL, &t, &n) && t < best)
		{
			best = t;
			bestN = n;
		}
	}
	if (best == FLT_MAX)
	{
		return false;
	}
	*fraction = best;
	*normal = m3d_rotate(xf.q, bestN);
	return true;
}

// Tight shape AABB; src/grid.cpp:145-174
// signed distance and closest surface point; src/world.cpp:1218-1284
MX_HD float point_shape_closest(const Shape& s, m3d_transform xf, m3d_vec3 p, m3d_vec3* closest)
{
	if (s.type == M3D_SHAPE_SPHERE)
	{
		m3d_vec3 d = m3d_sub(p, xf.p);
		float len = m3d_length(d);
		m3d_vec3 dir = len > 1e-9f ? m3d_scale(1.0f / len, d) : m3d_v3(0.0f, 1.0f, 0.0f);
		*closest = m3d_mul_add(xf.p, s.radius, dir);
		return len - s.radius;
	}
	if (s.type == M3D_SHAPE_CAPSULE)
	{
		m3d_vec3 axis = m3d_rotate(xf.q, m3d_v3(0.0f, 1.0f, 0.0f));
		m3d_vec3 pa = m3d_mul_add(xf.p, -s.halfHeight, axis);
		m3d_vec3 ab = m3d_scale(2.0f * s.halfHeight, axis);
		float den = m3d_dot(ab, ab);
		float t = den > 1e-12f ? m3d_dot(m3d_sub(p, pa), ab) / den : 0.0f;
		t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
		m3d_vec3 c = m3d_mul_add(pa, t, ab);
		m3d_vec3 d = m3d_sub(p, c);
		float len = m3d_length(d);
		m3d_vec3 dir = len > 1e-9f ? m3d_scale(1.0f / len, d) : m3d_v3(0.0f, 1.0f, 0.0f);
		*closest = m3d_mul_add(c, s.radius, dir);
		return len - s.radius;
	}
	m3d_vec3 lp = m3d_inv_transform_point(xf, p);
	m3d_vec3 he = s.halfExtents;
	m3d_vec3 cl = clamp_to_box(lp, he);
	m3d_vec3 diff = m3d_sub(lp, cl);
	float d2 = m3d_length_sq(diff);
	if (d2 > 0.0f)
	{
		*closest = m3d_transform_point(xf, cl);
		return sqrtf(d2);
	}
	float dx = he.x - fabsf(lp.x);
	float dy = he.y - fabsf(lp.y);
	float dz = he.z - fabsf(lp.z);
	m3d_vec3 face = lp;
	float depth;
	if (dx <= dy && dx <= dz)
	{
		face.x = lp.x >= 0.0f ? he.x : -he.x;
		depth = dx;
	}
	else if (dy <= dz)
	{
		face.y = lp.y >= 0.0f ? he.y : -he.y;
		depth = dy;
	}
	else
	{
		face.z = lp.z >= 0.0f ? he.z : -he.z;
		depth = dz;
	}
	*closest = m3d_transform_point(xf, face);
	return -depth;
}

MX_HD m3d_aabb shape_aabb(const Shape& s, m3d_transform xf)
{
	m3d_aabb r;
	if (s.type == M3D_SHAPE_SPHERE)
	{
		m3d_vec3 rr = m3d_v3(s.radius, s.radius, s.radius);
		r.lowerBound = m3d_sub(xf.p, rr);
		r.upperBound = m3d_add(xf.p, rr);
		return r;
	}
	if (s.type == M3D_SHAPE_CAPSULE)
	{
		m3d_vec3 ax = m3d_rotate(xf.q, m3d_v3(0.0f, s.halfHeight, 0.0f));
		m3d_vec3 p1 = m3d_add(xf.p, ax);
		m3d_vec3 p2 = m3d_sub(xf.p, ax);
		m3d_vec3 rr = m3d_v3(s.radius, s.radius, s.radius);
		r.lowerBound = m3d_sub(m3d_min(p1, p2), rr);
		r.upperBound = m3d_add(m3d_max(p1, p2), rr);
		return r;
	}
	m3d_mat3 rot = m3d_quat_to_mat3(xf.q);
	m3d_vec3 e = s.halfExtents;
	m3d_vec3 ext;
	ext.x = fabsf(rot.cx.x) * e.x + fabsf(rot.cy.x) * e.y + fabsf(rot.cz.x) * e.z;
	ext.y = fabsf(rot.cx.y) * e.x + fabsf(rot.cy.y) * e.y + fabsf(rot.cz.y) * e.z;
	ext.z = fabsf(rot.cx.z) * e.x + fabsf(rot.cy.z) * e.y + fabsf(rot.cz.z) * e.z;
	r.lowerBound = m3d_sub(xf.p, ext);
	r.upperBound = m3d_add(xf.p, ext);
	return r;
}

} // namespace mx
