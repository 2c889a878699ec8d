// XPBD substep stages (host+device): body integration and velocity
// reconstruction, per-substep contact preparation with the safety
// re-collision and tunneling rescue, the normal projection in its two
// arithmetic forms (serial-order and 8-lane packet form), stiction friction,
// the velocity pass, and joints.
//
// Each function reproduces the floating-point expressions of the reference
// routine cited beside it; the two normal-projection forms differ in their
// rounding (division vs reciprocal, association of the separation sum) and
// each solve path uses the form its reference path uses.
#pragma once

#include "mx_collide.h"

namespace mx
{

// Single-precision transcendentals evaluated in double precision and rounded
// once: the correctly rounded float result (up to double-rounding ties), the
// same on host and device, so parity with the reference does not hinge on a
// library's last-ulp choices.
MX_HD float sin_cr(float x) { return (float)sin((double)x); }
MX_HD float cos_cr(float x) { return (float)cos((double)x); }
MX_HD float atan2_cr(float y, float x) { return (float)atan2((double)y, (double)x); }

// Working state of one body for the solver stages.
struct WB
{
	m3d_vec3 p;
	m3d_quat q;
	m3d_vec3 v;
	m3d_vec3 w;
	float invMass;
	m3d_mat3 iw;
};

// src/xpbd.cpp:21-25
MX_HD m3d_quat renorm_fast(m3d_quat q)
{
	float s = 1.5f - 0.5f * (q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
	m3d_quat r = { q.x * s, q.y * s, q.z * s, q.w * s };
	return r;
}

// q + 0.5 * (dw,0) * q, first-order renormalized (shared by the position
// corrections of src/xpbd.cpp:28-58 and :440-475)
MX_HD m3d_quat rotate_by(m3d_quat q, m3d_vec3 dw)
{
	m3d_quat wq = { dw.x, dw.y, dw.z, 0.0f };
	m3d_quat dq = m3d_quat_mul(wq, q);
	m3d_quat r = { q.x + 0.5f * dq.x, q.y + 0.5f * dq.y, q.z + 0.5f * dq.z, q.w + 0.5f * dq.w };
	return renorm_fast(r);
}

// src/xpbd.cpp:28-41
MX_HD void apply_position_correction(WB& b, m3d_vec3 p, m3d_vec3 r)
{
	if (b.invMass == 0.0f)
	{
		return;
	}
	b.p = m3d_mul_add(b.p, b.invMass, p);
	m3d_vec3 dw = m3d_mat3_mulv(&b.iw, m3d_cross(r, p));
	b.q = rotate_by(b.q, dw);
}

// src/xpbd.cpp:45-58
MX_HD void apply_axis_correction(m3d_vec3& pos, m3d_quat& rot, float invMass, float dl, m3d_vec3 n, m3d_vec3 angJ)
{
	if (invMass == 0.0f)
	{
		return;
	}
	pos = m3d_mul_add(pos, dl * invMass, n);
	rot = rotate_by(rot, m3d_scale(dl, angJ));
}

// src/xpbd.cpp:61-69
MX_HD void apply_axis_velocity(WB& b, float j, m3d_vec3 n, m3d_vec3 angJ)
{
	if (b.invMass == 0.0f)
	{
		return;
	}
	b.v = m3d_mul_add(b.v, j * b.invMass, n);
	b.w = m3d_mul_add(b.w, j, angJ);
}

// src/xpbd.cpp:71-82
MX_HD void apply_velocity_impulse(WB& b, m3d_vec3 p, m3d_vec3 r)
{
	if (b.invMass == 0.0f)
	{
		return;
	}
	b.v = m3d_mul_add(b.v, b.invMass, p);
	b.w = m3d_add(b.w, m3d_mat3_mulv(&b.iw, m3d_cross(r, p)));
}

// src/xpbd.cpp:85-89
MX_HD float generalized_inv_mass(const WB& b, m3d_vec3 r, m3d_vec3 axis)
{
	m3d_vec3 rn = m3d_cross(r, axis);
	return b.invMass + m3d_dot(rn, m3d_mat3_mulv(&b.iw, rn));
}

// Contact arm: radius along the normal for spheres, clamped axial
// coordinate plus radius for capsules, rotated anchor for boxes;
// src/xpbd.cpp:100-116
MX_HD m3d_vec3 contact_arm(int shapeType, float radius, float halfHeight, m3d_quat rot, m3d_vec3 local, m3d_vec3 n,
						   float sign)
{
	if (shapeType == M3D_SHAPE_SPHERE)
	{
		return m3d_scale(sign * radius, n);
	}
	if (shapeType == M3D_SHAPE_CAPSULE)
	{
		float ay = local.y < -halfHeight ? -halfHeight : (local.y > halfHeight ? halfHeight : local.y);
		m3d_vec3 axis = m3d_rotate(rot, m3d_v3(0.0f, 1.0f, 0.0f));
		return m3d_add(m3d_scale(ay, axis), m3d_scale(sign * radius, n));
	}
	return m3d_rotate(rot, local);
}

// I^-1_world = R diag(I^-1_local) R^T; src/m3d_internal.h:126-146
MX_HD m3d_mat3 inv_inertia_world(m3d_quat rot, m3d_vec3 d)
{
	m3d_mat3 r = m3d_quat_to_mat3(rot);
	m3d_mat3 rd;
	rd.cx = m3d_scale(d.x, r.cx);
	rd.cy = m3d_scale(d.y, r.cy);
	rd.cz = m3d_scale(d.z, r.cz);
	m3d_mat3 m;
	m.cx = m3d_v3(rd.cx.x * r.cx.x + rd.cy.x * r.cy.x + rd.cz.x * r.cz.x,
				  rd.cx.y * r.cx.x + rd.cy.y * r.cy.x + rd.cz.y * r.cz.x,
				  rd.cx.z * r.cx.x + rd.cy.z * r.cy.x + rd.cz.z * r.cz.x);
	m.cy = m3d_v3(rd.cx.x * r.cx.y + rd.cy.x * r.cy.y + rd.cz.x * r.cz.y,
				  rd.cx.y * r.cx.y + rd.cy.y * r.cy.y + rd.cz.y * r.cz.y,
				  rd.cx.z * r.cx.y + rd.cy.z * r.cy.y + rd.cz.z * r.cz.y);
	m.cz = m3d_v3(rd.cx.x * r.cx.z + rd.cy.x * r.cy.z + rd.cz.x * r.cz.z,
				  rd.cx.y * r.cx.z + rd.cy.y * r.cy.z + rd.cz.y * r.cz.z,
				  rd.cx.z * r.cx.z + rd.cy.z * r.cy.z + rd.cz.z * r.cz.z);
	return m;
}

// Anchor-tracked minimum separation at the current poses;
// src/m3d_internal.h:251-268
template <class M>
MX_HD float tracked_min_sep(m3d_vec3 pa, m3d_quat qa, m3d_vec3 pb, m3d_quat qb, const M& m)
{
	if (m.pointCount == 0)
	{
		return 1e30f;
	}
	float mn = 1e30f;
	for (int i = 0; i < m.pointCount; ++i)
	{
		const MPoint mp = m.point(i);
		m3d_vec3 rA = m3d_rotate(qa, mp.localAnchorA);
		m3d_vec3 rB = m3d_rotate(qb, mp.localAnchorB);
		float sep = mp.separationOffset + m3d_dot(m3d_sub(m3d_add(pb, rB), m3d_add(pa, rA)), m.normal);
		mn = sep < mn ? sep : mn;
	}
	return mn;
}

// Pose moved beyond 1 mm / ~1 mrad; src/m3d_internal.h:231-241
MX_HD bool pose_changed(m3d_vec3 p, m3d_quat q, m3d_vec3 cp, m3d_quat cq)
{
	const float tolPos2 = 1e-3f * 1e-3f;
	const float tolHalfAng2 = 0.25f * 1e-6f;
	if (m3d_length_sq(m3d_sub(p, cp)) > tolPos2)
	{
		return true;
	}
	m3d_quat d = m3d_quat_mul(q, m3d_quat_conj(cq));
	return d.x * d.x + d.y * d.y + d.z * d.z > tolHalfAng2;
}

// separationOffset baseline at the current poses (src/world.cpp:1792-1799,
// src/xpbd.cpp:836-844)
MX_HD void set_separation_offsets(Manifold& m, m3d_vec3 pa, m3d_quat qa, m3d_vec3 pb, m3d_quat qb)
{
	for (int i = 0; i < m.pointCount; ++i)
	{
		MPoint mp = m.point(i);
		m3d_vec3 rA = m3d_rotate(qa, mp.localAnchorA);
		m3d_vec3 rB = m3d_rotate(qb, mp.localAnchorB);
		float ref = m3d_dot(m3d_sub(m3d_add(pb, rB), m3d_add(pa, rA)), m.normal);
		mp.separationOffset = mp.separation - ref;
		m.set_point(i, mp);
	}
}

// ------------------------------------------------------------ integrate

struct IntegrateIn
{
	float h;
	m3d_vec3 gravity;
};

// One body's substep integration; src/xpbd.cpp:587-705. Reads and writes the
// body columns directly (dynamic bodies only).
MX_HD void integrate_body(const DBodies& B, int32_t i, float h, m3d_vec3 gravity)
{
	uint32_t fl = B.flags[i];
	if (bf_type(fl) != M3D_BODY_DYNAMIC)
	{
		return;
	}
	float4 P = B.pos[i], Q = B.rot[i], V = B.vel[i], W = B.angVel[i];
	float4 IL = B.invIL[i], I0 = B.iw0[i], I1 = B.iw1[i], I2 = B.iw2[i];
	m3d_vec3 pos = f4_v3(P);
	m3d_quat rot = f4_q(Q);
	m3d_vec3 v = f4_v3(V);
	m3d_vec3 w = f4_v3(W);
	float invMass = P.w;
	float gravityScale = V.w;
	float linDamp = IL.w;
	float angDamp = I0.w;
	m3d_vec3 il = f4_v3(IL);
	m3d_mat3 iw;
	iw.cx = f4_v3(I0);
	iw.cy = f4_v3(I1);
	iw.cz = f4_v3(I2);

	B.prevVel[i] = v3_f4(v, 0.0f);
	B.prevAngVel[i] = v3_f4(w, 0.0f);

	v = m3d_mul_add(v, h * gravityScale, gravity);
	m3d_vec3 f = f4_v3(B.force[i]);
	if (f.x != 0.0f || f.y != 0.0f || f.z != 0.0f)
	{
		v = m3d_mul_add(v, h * invMass, f);
	}
	m3d_vec3 t = f4_v3(B.torque[i]);
	if (t.x != 0.0f || t.y != 0.0f || t.z != 0.0f)
	{
		w = m3d_add(w, m3d_scale(h, m3d_mat3_mulv(&iw, t)));
	}
	v = m3d_scale(1.0f / (1.0f + h * linDamp), v);
	w = m3d_scale(1.0f / (1.0f + h * angDamp), w);

	// implicit gyroscopic term, one Newton step in the body frame
	bool aniso = il.x != il.y || il.y != il.z;
	if (aniso && il.x > 0.0f && il.y > 0.0f && il.z > 0.0f && m3d_length_sq(w) > 1.0f)
	{
		m3d_vec3 w0 = m3d_inv_rotate(rot, w);
		m3d_vec3 wl = w0;
		m3d_vec3 I = m3d_v3(1.0f / il.x, 1.0f / il.y, 1.0f / il.z);
		m3d_vec3 Iw = m3d_v3(I.x * wl.x, I.y * wl.y, I.z * wl.z);
		m3d_vec3 f0 = m3d_add(m3d_v3(I.x * (wl.x - w0.x), I.y * (wl.y - w0.y), I.z * (wl.z - w0.z)),
							  m3d_scale(h, m3d_cross(wl, Iw)));
		m3d_vec3 ex = m3d_v3(1, 0, 0), ey = m3d_v3(0, 1, 0), ez = m3d_v3(0, 0, 1);
		m3d_vec3 c0 = m3d_add(m3d_v3(I.x, 0, 0),
							  m3d_scale(h, m3d_sub(m3d_scale(I.x, m3d_cross(wl, ex)), m3d_cross(Iw, ex))));
		m3d_vec3 c1 = m3d_add(m3d_v3(0, I.y, 0),
							  m3d_scale(h, m3d_sub(m3d_scale(I.y, m3d_cross(wl, ey)), m3d_cross(Iw, ey))));
		m3d_vec3 c2 = m3d_add(m3d_v3(0, 0, I.z),
							  m3d_scale(h, m3d_sub(m3d_scale(I.z, m3d_cross(wl, ez)), m3d_cross(Iw, ez))));
		float det = m3d_dot(c0, m3d_cross(c1, c2));
		if (!(fabsf(det) <= 1e-12f))
		{
			m3d_vec3 rhs = m3d_neg(f0);
			float invDet = 1.0f / det;
			m3d_vec3 dw = m3d_v3(m3d_dot(rhs, m3d_cross(c1, c2)) * invDet, m3d_dot(c0, m3d_cross(rhs, c2)) * invDet,
								 m3d_dot(c0, m3d_cross(c1, rhs)) * invDet);
			wl = m3d_add(wl, dw);
		}
		w = m3d_rotate(rot, wl);
	}

	B.prevPos[i] = v3_f4(pos, 0.0f);
	B.prevRot[i] = q_f4(rot);
	pos = m3d_mul_add(pos, h, v);

	// exponential-map rotation update
	float wlen = m3d_length(w);
	if (wlen > 1e-9f)
	{
		float half = 0.5f * h * wlen;
		float sh, ch;
		if (half < 0.02f)
		{
			float h2 = half * half;
			sh = half * (1.0f - h2 * (1.0f / 6.0f));
			ch = 1.0f - h2 * 0.5f;
		}
		else
		{
			sh = sin_cr(half);
			ch = cos_cr(half);
		}
		float s = sh / wlen;
		m3d_quat dq = { w.x * s, w.y * s, w.z * s, ch };
		rot = renorm_fast(m3d_quat_mul(dq, rot));
	}
	iw = inv_inertia_world(rot, il);

	B.pos[i] = v3_f4(pos, invMass);
	B.rot[i] = q_f4(rot);
	B.vel[i] = v3_f4(v, gravityScale);
	B.angVel[i] = v3_f4(w, W.w);
	B.iw0[i] = v3_f4(iw.cx, I0.w);
	B.iw1[i] = v3_f4(iw.cy, I1.w);
	B.iw2[i] = v3_f4(iw.cz, I2.w);
}

// src/xpbd.cpp:865-892
MX_HD void reconstruct_velocity(const DBodies& B, int32_t i, float invH)
{
	if (bf_type(B.flags[i]) != M3D_BODY_DYNAMIC)
	{
		return;
	}
	float4 P = B.pos[i];
	m3d_vec3 pos = f4_v3(P);
	m3d_quat rot = f4_q(B.rot[i]);
	m3d_vec3 pp = f4_v3(B.prevPos[i]);
	m3d_quat pq = f4_q(B.prevRot[i]);
	float gs = B.vel[i].w;
	float mass = B.angVel[i].w;
	m3d_vec3 v = m3d_scale(invH, m3d_sub(pos, pp));
	m3d_quat dq = m3d_quat_mul(rot, m3d_quat_conj(pq));
	float sign = dq.w >= 0.0f ? 1.0f : -1.0f;
	float v2 = dq.x * dq.x + dq.y * dq.y + dq.z * dq.z;
	m3d_vec3 w;
	if (v2 > 4e-4f)
	{
		float vlen = sqrtf(v2);
		float angle = 2.0f * atan2_cr(vlen, fabsf(dq.w));
		w = m3d_scale(sign * invH * angle / vlen, m3d_v3(dq.x, dq.y, dq.z));
	}
	else
	{
		float corr = 1.0f + v2 * (1.0f / 6.0f);
		w = m3d_scale(2.0f * invH * sign * corr, m3d_v3(dq.x, dq.y, dq.z));
	}
	B.vel[i] = v3_f4(v, gs);
	B.angVel[i] = v3_f4(w, mass);
}

// ------------------------------------------------------------- prepare

// Safety re-collision trigger: > 3 cm or ~5.7 deg since the cached pose;
// src/xpbd.cpp:718-731
MX_HD bool manifold_stale(m3d_vec3 pa, m3d_quat qa, m3d_vec3 pb, m3d_quat qb, m3d_vec3 cpa, m3d_quat cqa, m3d_vec3 cpb,
						  m3d_quat cqb)
{
	const float pos2 = 0.03f * 0.03f;
	const float ang2 = 0.05f * 0.05f;
	if (m3d_length_sq(m3d_sub(pa, cpa)) > pos2 || m3d_length_sq(m3d_sub(pb, cpb)) > pos2)
	{
		return true;
	}
	m3d_quat da = m3d_quat_mul(qa, m3d_quat_conj(cqa));
	m3d_quat db = m3d_quat_mul(qb, m3d_quat_conj(cqb));
	return da.x * da.x + da.y * da.y + da.z * da.z > ang2 || db.x * db.x + db.y * db.y + db.z * db.z > ang2;
}

MX_HD float min3(float a, float b, float c) { return fminf(a, fminf(b, c)); }

// Per-substep contact preparation; src/xpbd.cpp:733-863.
// Receives each point's prepared rows (anchor A, anchor B with wN, angular
// Jacobians A and B, arms A and B) as they are stored, after a preload call
// per point made with the contact's other loads, and the contact's normal and
// material words as they stand at the end; the default drops them.
struct NoPrepSink
{
	MX_HD void preload(int, int32_t) const {}
	MX_HD void point(int, int32_t, const float4&, const float4&, const float4&, const float4&, const float4&,
					 const float4&) const
	{
	}
	MX_HD void contact(const float4&, const float4&) const {}
};

template <class Sink = NoPrepSink>
MX_HD void prepare_contact(const DBodies& B, const DContacts& C, int32_t c, float h, const Sink& sink = Sink())
{
	int2 ab = C.bodies[c];
	int32_t ia = ab.x, ib = ab.y;
	float4 mat = C.mat[c];
	uint32_t cf = f2u(mat.w);
	float4 PA = B.pos[ia], PB = B.pos[ib];
	m3d_vec3 pa = f4_v3(PA), pb = f4_v3(PB);
	m3d_quat qa = f4_q(B.rot[ia]), qb = f4_q(B.rot[ib]);
	float imA = PA.w, imB = PB.w;

	float4 nc = C.normal[c];
	m3d_vec3 n = f4_v3(nc);
	int32_t count = f2i(nc.w);

	bool hasCache = (cf & CF_HAS_CACHE) != 0;
	bool stale = !hasCache;
	if (hasCache)
	{
		const float4* k = cache_of(C, c);
		stale = manifold_stale(pa, qa, pb, qb, f4_v3(k[0]), f4_q(k[1]), f4_v3(k[2]), f4_q(k[3]));
	}
	if (stale)
	{
		// the old manifold is read where it lies; the new one is built here
		// (slots beyond its count are never stored, so they keep their old
		// contents, as the reference's in-place regeneration leaves them)
		const StoredManifold old(C, c);
		Manifold m;
		Shape sA = load_shape(B, ia), sB = load_shape(B, ib);
		m3d_transform xfA = { pa, qa }, xfB = { pb, qb };
		float relSpeed = m3d_length(m3d_sub(f4_v3(B.vel[ia]), f4_v3(B.vel[ib])));
		collide_shapes(sA, xfA, sB, xfB, kSpeculativeBase + h * relSpeed, &m);

		bool keepOld = false;
		if (hasCache && old.pointCount > 0)
		{
			if (cf & CF_TUNNEL_GUARD)
			{
				keepOld = tracked_min_sep(pa, qa, pb, qb, old) < -kLinearSlop;
			}
			else
			{
				bool flipOrGone = m.pointCount == 0 || m3d_dot(m.normal, old.normal) < -0.5f;
				if (flipOrGone)
				{
					int ta = bf_type(B.flags[ia]), tb = bf_type(B.flags[ib]);
					m3d_vec3 dB = tb == M3D_BODY_DYNAMIC ? m3d_sub(pb, f4_v3(B.prevPos[ib])) : m3d_v3_zero();
					m3d_vec3 dA = ta == M3D_BODY_DYNAMIC ? m3d_sub(pa, f4_v3(B.prevPos[ia])) : m3d_v3_zero();
					m3d_vec3 travel = m3d_sub(dB, dA);
					float travelLen = m3d_length(travel);
					float coreB = sB.type == M3D_SHAPE_BOX
									  ? min3(sB.halfExtents.x, sB.halfExtents.y, sB.halfExtents.z)
									  : sB.radius;
					float minHalfA = sA.type == M3D_SHAPE_BOX
										 ? min3(sA.halfExtents.x, sA.halfExtents.y, sA.halfExtents.z)
										 : sA.radius;
					float crossSpan = 2.0f * (minHalfA + coreB);
					if (travelLen > 0.04f && travelLen > 0.8f * crossSpan)
					{
						Shape inflated = sA;
						if (inflated.type == M3D_SHAPE_BOX)
						{
							inflated.halfExtents = m3d_add(inflated.halfExtents, m3d_v3(coreB, coreB, coreB));
						}
						else
						{
							inflated.radius += coreB;
						}
						float backup = 2.0f * coreB + 0.02f;
						m3d_vec3 dir = m3d_scale(1.0f / travelLen, travel);
						m3d_vec3 origin = m3d_sub(m3d_sub(pb, dB), m3d_scale(backup, dir));
						m3d_vec3 seg = m3d_add(travel, m3d_scale(backup, dir));
						float fh;
						m3d_vec3 nh;
						keepOld = ray_cast_shape(inflated, xfA, origin, seg, &fh, &nh);
					}
				}
			}
			cf = keepOld ? (cf | CF_TUNNEL_GUARD) : (cf & ~CF_TUNNEL_GUARD);
		}
		if (keepOld)
		{
			// the old manifold stays as stored (cache pose intentionally
			// not refreshed)
			n = old.normal;
			count = old.pointCount;
		}
		else
		{
			inherit_friction_anchors(old, m, pa, qa, pb, qb,
									 sA.type == M3D_SHAPE_BOX && sB.type == M3D_SHAPE_BOX);
			cf |= CF_HAS_CACHE;
			float4* k = cache_of(C, c);
			k[0] = v3_f4(pa, 0.0f);
			k[1] = q_f4(qa);
			k[2] = v3_f4(pb, 0.0f);
			k[3] = q_f4(qb);
			set_separation_offsets(m, pa, qa, pb, qb);
			store_manifold(C, c, m);
			n = m.normal;
			count = m.pointCount;
		}
		C.mat[c] = make_float4(mat.x, mat.y, mat.z, u2f(cf));
	}

	// freeze arms, Jacobians and normal mass for this substep. Every point's
	// inputs are loaded before the first store: on this device a load issued
	// after a store is waited for together with the store's completion, so
	// loads between the points' stores would cost a memory round trip each.
	m3d_mat3 iwA = load_iw(B, ia), iwB = load_iw(B, ib);
	float4 SA = B.shape[ia], SB = B.shape[ib];
	int tA = bf_shape(B.flags[ia]), tB = bf_shape(B.flags[ib]);
	float hhA = B.mat[ia].x, hhB = B.mat[ib].x;
	float4 a4v[4], b4v[4], rav[4], rbv[4];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		if (j < count)
		{
			int32_t p = pidx(C, c, j);
			a4v[j] = C.pA[p];
			b4v[j] = C.pB[p];
			rav[j] = C.pRA[p];
			rbv[j] = C.pRB[p];
			sink.preload(j, p);
		}
	}
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		if (j < count)
		{
			int32_t p = pidx(C, c, j);
			float4 a4 = a4v[j], b4 = b4v[j];
			m3d_vec3 rA = contact_arm(tA, SA.w, hhA, qa, f4_v3(a4), n, 1.0f);
			m3d_vec3 rB = contact_arm(tB, SB.w, hhB, qb, f4_v3(b4), n, -1.0f);
			m3d_vec3 crA = m3d_cross(rA, n);
			m3d_vec3 crB = m3d_cross(rB, n);
			m3d_vec3 jA = m3d_mat3_mulv(&iwA, crA);
			m3d_vec3 jB = m3d_mat3_mulv(&iwB, crB);
			float wN = imA + imB + m3d_dot(crA, jA) + m3d_dot(crB, jB);
			const float4 bw = make_float4(b4.x, b4.y, b4.z, wN);
			const float4 ja = v3_f4(jA, 0.0f); // lambdaN = 0
			const float4 jb = v3_f4(jB, 0.0f); // lambdaT = 0
			C.pB[p] = bw;
			C.pJA[p] = ja;
			C.pJB[p] = jb;
			float4 ra = rav[j], rb = rbv[j];
			const float4 ra2 = v3_f4(rA, ra.w), rb2 = v3_f4(rB, rb.w);
			C.pRA[p] = ra2;
			C.pRB[p] = rb2;
			sink.point(j, p, a4, bw, ja, jb, ra2, rb2);
		}
	}
	if (stale)
	{
		// as the regeneration (or the kept manifold) left them
		nc = C.normal[c];
		mat = C.mat[c];
	}
	sink.contact(nc, mat);
}

// ------------------------------------------------------ normal projection

// Contact-side inputs of a normal projection. Every point slot is allocated
// (4 per contact), so all of them are loaded at once, before the point count
// is known: one memory round trip, which a caller may also issue early (only
// the contact itself writes these values).
struct NormalsIn
{
	float4 nc;
	float4 A4[4], B4[4], JA[4], JB[4];
};

MX_HD void load_normals_in(const DContacts& C, int32_t c, NormalsIn& d)
{
	d.nc = C.normal[c];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = pidx(C, c, j);
		d.A4[j] = C.pA[p];
		d.B4[j] = C.pB[p];
		d.JA[j] = C.pJA[p];
		d.JB[j] = C.pJB[p];
	}
}

// Serial-order form (SolveContactNormals, src/xpbd.cpp:122-163): divides by
// wN and skips points with wN <= 0 or sep >= 0. Updates the body states in
// place and accumulates lambdaN.
MX_HD void solve_normals_serial_in(const DContacts& C, int32_t c, const NormalsIn& d, m3d_vec3& pa, m3d_quat& qa,
								   float imA, m3d_vec3& pb, m3d_quat& qb, float imB, float maxCorrection)
{
	const float4 nc = d.nc;
	const float4* A4 = d.A4;
	const float4* B4 = d.B4;
	const float4* JA = d.JA;
	const float4* JB = d.JB;
	m3d_vec3 n = f4_v3(nc);
	int32_t count = f2i(nc.w);
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		if (j >= count)
		{
			break;
		}
		float4 b4 = B4[j];
		float wN = b4.w;
		if (wN <= 0.0f)
		{
			continue;
		}
		float4 a4 = A4[j];
		m3d_vec3 rA = m3d_rotate(qa, f4_v3(a4));
		m3d_vec3 rB = m3d_rotate(qb, f4_v3(b4));
		float sep = a4.w + m3d_dot(m3d_sub(m3d_add(pb, rB), m3d_add(pa, rA)), n);
		if (sep >= 0.0f)
		{
			continue;
		}
		if (sep < -maxCorrection)
		{
			sep = -maxCorrection;
		}
		float dl = -sep / wN;
		float4 ja = JA[j], jb = JB[j];
		C.pJA[pidx(C, c, j)].w = ja.w + dl;
		apply_axis_correction(pb, qb, imB, dl, n, f4_v3(jb));
		apply_axis_correction(pa, qa, imA, -dl, n, f4_v3(ja));
	}
}

MX_HD void solve_normals_serial(const DContacts& C, int32_t c, m3d_vec3& pa, m3d_quat& qa, float imA, m3d_vec3& pb,
								m3d_quat& qb, float imB, float maxCorrection)
{
	NormalsIn d;
	load_normals_in(C, c, d);
	solve_normals_serial_in(C, c, d, pa, qa, imA, pb, qb, imB, maxCorrection);
}

// 8-lane packet form (SolvePacketNormals, src/xpbd.cpp:1047-1141) applied to
// one lane: points iterate to the packet's maxPoints; padding points use a
// zero point with invW = 1 and a zero gate; both bodies' lane copies are
// updated every point (the caller writes back dynamic bodies only).
MX_HD void solve_normals_packet_in(const DContacts& C, int32_t c, const NormalsIn& d, int32_t maxPoints,
								   m3d_vec3& pa, m3d_quat& qa, float imA, m3d_vec3& pb, m3d_quat& qb, float imB,
								   float maxCorrection)
{
	const float4 nc = d.nc;
	const float4* A4 = d.A4;
	const float4* B4 = d.B4;
	const float4* JA = d.JA;
	const float4* JB = d.JB;
	float nx = nc.x, ny = nc.y, nz = nc.z;
	int32_t count = f2i(nc.w);
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		if (j >= maxPoints)
		{
			break;
		}
		bool on = j < count;
		int32_t p = pidx(C, c, on ? j : 0);
		const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		float4 a4 = on ? A4[j] : zero;
		float4 b4 = on ? B4[j] : zero;
		float4 ja = on ? JA[j] : zero;
		float4 jb = on ? JB[j] : zero;
		float active = on ? 1.0f : 0.0f;
		float invW = (on && b4.w > 0.0f) ? 1.0f / b4.w : 1.0f;

		float vx = a4.x, vy = a4.y, vz = a4.z;
		float tx = 2.0f * (qa.y * vz - qa.z * vy);
		float ty = 2.0f * (qa.z * vx - qa.x * vz);
		float tz = 2.0f * (qa.x * vy - qa.y * vx);
		float rax = vx + qa.w * tx + (qa.y * tz - qa.z * ty);
		float ray = vy + qa.w * ty + (qa.z * tx - qa.x * tz);
		float raz = vz + qa.w * tz + (qa.x * ty - qa.y * tx);
		vx = b4.x;
		vy = b4.y;
		vz = b4.z;
		tx = 2.0f * (qb.y * vz - qb.z * vy);
		ty = 2.0f * (qb.z * vx - qb.x * vz);
		tz = 2.0f * (qb.x * vy - qb.y * vx);
		float rbx = vx + qb.w * tx + (qb.y * tz - qb.z * ty);
		float rby = vy + qb.w * ty + (qb.z * tx - qb.x * tz);
		float rbz = vz + qb.w * tz + (qb.x * ty - qb.y * tx);

		float sep = a4.w + (pb.x + rbx - pa.x - rax) * nx + (pb.y + rby - pa.y - ray) * ny +
					(pb.z + rbz - pa.z - raz) * nz;
		float gate = sep < 0.0f ? active : 0.0f;
		float capped = sep < -maxCorrection ? -maxCorrection : sep;
		float dl = -capped * invW * gate;
		if (on)
		{
			C.pJA[p].w = ja.w + dl;
		}

		pb.x += dl * imB * nx;
		pb.y += dl * imB * ny;
		pb.z += dl * imB * nz;
		pa.x -= dl * imA * nx;
		pa.y -= dl * imA * ny;
		pa.z -= dl * imA * nz;

		float dwx = dl * jb.x, dwy = dl * jb.y, dwz = dl * jb.z;
		float dqx = dwx * qb.w + dwy * qb.z - dwz * qb.y;
		float dqy = -dwx * qb.z + dwy * qb.w + dwz * qb.x;
		float dqz = dwx * qb.y - dwy * qb.x + dwz * qb.w;
		float dqw = -(dwx * qb.x + dwy * qb.y + dwz * qb.z);
		float x = qb.x + 0.5f * dqx, y = qb.y + 0.5f * dqy, z = qb.z + 0.5f * dqz, w = qb.w + 0.5f * dqw;
		float s = 1.5f - 0.5f * (x * x + y * y + z * z + w * w);
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
		x = qa.x + 0.5f * dqx;
		y = qa.y + 0.5f * dqy;
		z = qa.z + 0.5f * dqz;
		w = qa.w + 0.5f * dqw;
		s = 1.5f - 0.5f * (x * x + y * y + z * z + w * w);
		qa.x = x * s;
		qa.y = y * s;
		qa.z = z * s;
		qa.w = w * s;
	}
}

MX_HD void solve_normals_packet(const DContacts& C, int32_t c, int32_t maxPoints, m3d_vec3& pa, m3d_quat& qa,
								float imA, m3d_vec3& pb, m3d_quat& qb, float imB, float maxCorrection)
{
	NormalsIn d;
	load_normals_in(C, c, d);
	solve_normals_packet_in(C, c, d, maxPoints, pa, qa, imA, pb, qb, imB, maxCorrection);
}

// -------------------------------------------------------------- friction

// Positional stiction toward persistent anchors, box-box only;
// src/xpbd.cpp:172-269
// Contact-side inputs of the stiction pass, loaded at once (every point
// slot is allocated).
struct FrictionIn
{
	float4 nc;
	float mu;
	float4 RA0, RB0;
	float LN[4];
	float4 FA[4], FB[4], JB[4];
};

MX_HD void load_friction_in(const DContacts& C, int32_t c, FrictionIn& d)
{
	d.nc = C.normal[c];
	d.mu = C.mat[c].x;
	int32_t p0 = pidx(C, c, 0);
	d.RA0 = C.pRA[p0];
	d.RB0 = C.pRB[p0];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = pidx(C, c, j);
		d.LN[j] = C.pJA[p].w;
		d.FA[j] = C.pFA[p];
		d.FB[j] = C.pFB[p];
		d.JB[j] = C.pJB[p];
	}
}

MX_HD void apply_friction_in(const DContacts& C, int32_t c, const FrictionIn& d, WB& a, WB& b)
{
	const float4 nc = d.nc;
	const float mu = d.mu;
	const float4 RA0 = d.RA0, RB0 = d.RB0;
	const float* LN = d.LN;
	const float4* FA = d.FA;
	const float4* FB = d.FB;
	const float4* JB = d.JB;
	int32_t count = f2i(nc.w);
	if (count == 0)
	{
		return;
	}
	m3d_vec3 n = f4_v3(nc);
	{
		const float maxSpeed2 = 0.1f * 0.1f;
		m3d_vec3 rA0 = f4_v3(RA0), rB0 = f4_v3(RB0);
		m3d_vec3 vrel = m3d_sub(m3d_add(b.v, m3d_cross(b.w, rB0)), m3d_add(a.v, m3d_cross(a.w, rA0)));
		m3d_vec3 vt = m3d_sub(vrel, m3d_scale(m3d_dot(vrel, n), n));
		if (m3d_length_sq(vt) > maxSpeed2)
		{
			for (int j = 0; j < count; ++j)
			{
				int32_t p = pidx(C, c, j);
				C.pFA[p] = C.pA[p];
				C.pFB[p] = C.pB[p];
			}
			return;
		}
	}
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		if (j >= count)
		{
			break;
		}
		int32_t p = pidx(C, c, j);
		float lambdaN = LN[j];
		if (lambdaN <= 0.0f)
		{
			continue;
		}
		m3d_vec3 rA = m3d_rotate(a.q, f4_v3(FA[j]));
		m3d_vec3 rB = m3d_rotate(b.q, f4_v3(FB[j]));
		m3d_vec3 gap = m3d_sub(m3d_add(b.p, rB), m3d_add(a.p, rA));
		m3d_vec3 dpt = m3d_sub(gap, m3d_scale(m3d_dot(gap, n), n));
		float lt = m3d_length(dpt);
		if (lt <= 1e-9f)
		{
			continue;
		}
		m3d_vec3 t = m3d_scale(1.0f / lt, dpt);
		float wt = generalized_inv_mass(a, rA, t) + generalized_inv_mass(b, rB, t);
		if (wt <= 0.0f)
		{
			continue;
		}
		float needed = kFrictionRelaxation * (lt / wt);
		float4 jb = JB[j];
		float lambdaT = jb.w;
		float allowed = mu * lambdaN - lambdaT;
		if (allowed <= 0.0f)
		{
			continue;
		}
		float dL = needed < allowed ? needed : allowed;
		C.pJB[p].w = lambdaT + dL;
		m3d_vec3 pt = m3d_scale(dL, t);
		apply_position_correction(b, m3d_neg(pt), rB);
		apply_position_correction(a, pt, rA);
		if (needed > 2.0f * allowed)
		{
			C.pFA[p] = C.pA[p];
			C.pFB[p] = C.pB[p];
		}
	}
}

MX_HD void apply_friction(const DContacts& C, int32_t c, bool boxBox, WB& a, WB& b)
{
	if (!boxBox)
	{
		return;
	}
	FrictionIn d;
	load_friction_in(C, c, d);
	apply_friction_in(C, c, d, a, b);
}

// ------------------------------------------------------------ velocity

struct VelShape
{
	int type;
	float radius;
	float halfHeight;
};

// Dynamic friction, restitution (applied unconditionally, XPBD eq. 29) and
// per-contact rolling resistance; src/xpbd.cpp:275-407.
// Contact-side inputs of the velocity pass, loaded at once: manifold,
// per-point impulses and slip, and the lever-arm sources (anchors for curved
// pairs, stored arms otherwise).
struct VelocityIn
{
	float4 nc, mat;
	float4 JA[4], B4[4], SL[4], JBv[4], AR[4], BR[4];
};

MX_HD void load_velocity_in(const DContacts& C, int32_t c, VelocityIn& d)
{
	d.nc = C.normal[c];
	d.mat = C.mat[c];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = pidx(C, c, j);
		d.JA[j] = C.pJA[p];
		d.B4[j] = C.pB[p];
		d.SL[j] = C.pSlip[p];
		d.JBv[j] = C.pJB[p];
	}
	bool curved = (f2u(d.mat.w) & CF_CURVED) != 0;
	const float4* srcA = curved ? C.pA : C.pRA;
	const float4* srcB = curved ? C.pB : C.pRB;
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		int32_t p = pidx(C, c, j);
		d.AR[j] = srcA[p];
		d.BR[j] = srcB[p];
	}
}

MX_HD void solve_velocity_in(const DContacts& C, int32_t c, const VelocityIn& d, float h, float gravityLen, WB& a,
							 WB& b, const m3d_vec3& pvA, const m3d_vec3& pwA, const m3d_vec3& pvB,
							 const m3d_vec3& pwB, const VelShape& shA, const VelShape& shB)
{
	const float4 nc = d.nc, mat = d.mat;
	const float4* JA = d.JA;
	const float4* B4 = d.B4;
	const float4* SL = d.SL;
	const float4* JBv = d.JBv;
	const float4* AR = d.AR;
	const float4* BR = d.BR;
	m3d_vec3 n = f4_v3(nc);
	int32_t count = f2i(nc.w);
	float mu = mat.x, e0 = mat.y, rr = mat.z;
	bool curved = (f2u(mat.w) & CF_CURVED) != 0;
	float invH = h > 0.0f ? 1.0f / h : 0.0f;
	float lambdaNTotal = 0.0f;

#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		if (j >= count)
		{
			break;
		}
		int32_t p = pidx(C, c, j);
		float4 ja = JA[j];
		float lambdaN = ja.w;
		if (lambdaN <= 0.0f)
		{
			continue;
		}
		float4 b4 = B4[j];
		m3d_vec3 rAf, rBf;
		if (curved)
		{
			rAf = contact_arm(shA.type, shA.radius, shA.halfHeight, a.q, f4_v3(AR[j]), n, 1.0f);
			rBf = contact_arm(shB.type, shB.radius, shB.halfHeight, b.q, f4_v3(BR[j]), n, -1.0f);
		}
		else
		{
			rAf = f4_v3(AR[j]);
			rBf = f4_v3(BR[j]);
		}
		m3d_vec3 vrel = m3d_sub(m3d_add(b.v, m3d_cross(b.w, rBf)), m3d_add(a.v, m3d_cross(a.w, rAf)));
		float vn = m3d_dot(vrel, n);
		m3d_vec3 vt = m3d_sub(vrel, m3d_scale(vn, n));
		float vtLen = m3d_length(vt);
		if (vtLen > 1e-9f)
		{
			m3d_vec3 t = m3d_scale(1.0f / vtLen, vt);
			float wt = generalized_inv_mass(a, rAf, t) + generalized_inv_mass(b, rBf, t);
			if (wt > 0.0f)
			{
				const float staticSlip = 0.08f;
				m3d_vec3 slip = f4_v3(SL[j]);
				bool coherent = m3d_dot(t, slip) > 0.7f;
				float pStop = vtLen / wt;
				float pMax = mu * lambdaN * invH;
				float P = !coherent && vtLen < staticSlip ? pStop : (pStop < pMax ? pStop : pMax);
				C.pSlip[p] = v3_f4(t, 0.0f);
				m3d_vec3 imp = m3d_scale(-P, t);
				apply_velocity_impulse(b, imp, rBf);
				apply_velocity_impulse(a, m3d_neg(imp), rAf);
			}
		}
		m3d_vec3 vrelPrev = m3d_sub(m3d_add(pvB, m3d_cross(pwB, rBf)), m3d_add(pvA, m3d_cross(pwA, rAf)));
		float vnPrev = m3d_dot(vrelPrev, n);
		float e = fabsf(vnPrev) > 2.0f * gravityLen * h ? e0 : 0.0f;
		vrel = m3d_sub(m3d_add(b.v, m3d_cross(b.w, rBf)), m3d_add(a.v, m3d_cross(a.w, rAf)));
		vn = m3d_dot(vrel, n);
		float target = -e * vnPrev;
		target = target > 0.0f ? target : 0.0f;
		float wN = b4.w;
		if (wN > 0.0f)
		{
			float jj = (target - vn) / wN;
			apply_axis_velocity(b, jj, n, f4_v3(JBv[j]));
			apply_axis_velocity(a, -jj, n, f4_v3(ja));
		}
		lambdaNTotal += lambdaN > 0.0f ? lambdaN : 0.0f;
	}

	if (rr > 0.0f && lambdaNTotal > 0.0f)
	{
		m3d_vec3 wrel = m3d_sub(b.w, a.w);
		m3d_vec3 wroll = m3d_sub(wrel, m3d_scale(m3d_dot(wrel, n), n));
		float wl = m3d_length(wroll);
		if (wl > 1e-6f)
		{
			m3d_vec3 axis = m3d_scale(1.0f / wl, wroll);
			float k = m3d_dot(axis, m3d_mat3_mulv(&a.iw, axis)) + m3d_dot(axis, m3d_mat3_mulv(&b.iw, axis));
			if (k > 1e-12f)
			{
				float maxDw = rr * lambdaNTotal * invH * k;
				float dw = maxDw < wl ? maxDw : wl;
				m3d_vec3 angImp = m3d_scale(dw / k, axis);
				b.w = m3d_sub(b.w, m3d_mat3_mulv(&b.iw, angImp));
				a.w = m3d_add(a.w, m3d_mat3_mulv(&a.iw, angImp));
			}
		}
	}
}

MX_HD void solve_velocity(const DContacts& C, int32_t c, float h, float gravityLen, WB& a, WB& b,
						  const m3d_vec3& pvA, const m3d_vec3& pwA, const m3d_vec3& pvB, const m3d_vec3& pwB,
						  const VelShape& shA, const VelShape& shB)
{
	VelocityIn d;
	load_velocity_in(C, c, d);
	solve_velocity_in(C, c, d, h, gravityLen, a, b, pvA, pwA, pvB, pwB, shA, shB);
}

// ---------------------------------------------------------------- joints

// Device copy of a joint (src/m3d_internal.h:313-349); type 0 distance,
// 1 ball, 2 hinge, 3 weld.
struct DJoint
{
	int32_t type;
	int32_t bodyA;
	int32_t bodyB;
	int32_t inUse;
	m3d_vec3 localAnchorA;
	m3d_vec3 localAnchorB;
	float length;
	float hertz;
	float dampingRatio;
	m3d_vec3 localAxisA;
	m3d_vec3 localAxisB;
	m3d_vec3 localRefA;
	m3d_vec3 localRefB;
	int32_t enableLimit;
	float lowerAngle;
	float upperAngle;
	int32_t enableMotor;
	float motorSpeed;
	float maxMotorTorque;
	m3d_quat relRot0;
	float compliance;
	float lambda;
	float motorLambda;
};

// src/m3d_internal.h:354-362
MX_HD float hinge_angle(m3d_quat qa, m3d_quat qb, const DJoint& j)
{
	m3d_vec3 axis = m3d_rotate(qa, j.localAxisA);
	m3d_vec3 refA = m3d_rotate(qa, j.localRefA);
	m3d_vec3 refB = m3d_rotate(qb, j.localRefB);
	float s = m3d_dot(axis, m3d_cross(refA, refB));
	float c = m3d_dot(refA, refB);
	return atan2_cr(s, c);
}

// src/xpbd.cpp:413-433
MX_HD void prepare_joint(DJoint& j, const WB& a, const WB& b)
{
	j.compliance = 0.0f;
	if (j.type == 0 && j.hertz > 0.0f)
	{
		m3d_vec3 rA = m3d_rotate(a.q, j.localAnchorA);
		m3d_vec3 rB = m3d_rotate(b.q, j.localAnchorB);
		m3d_vec3 d = m3d_sub(m3d_add(b.p, rB), m3d_add(a.p, rA));
		m3d_vec3 u = m3d_normalize(d);
		float w = generalized_inv_mass(a, rA, u) + generalized_inv_mass(b, rB, u);
		if (w > 0.0f)
		{
			float meff = 1.0f / w;
			float omega = 2.0f * M3D_PI * j.hertz;
			float k = meff * omega * omega;
			j.compliance = k > 0.0f ? 1.0f / k : 0.0f;
		}
	}
}

// src/xpbd.cpp:440-475
MX_HD void apply_angular_correction(WB& a, WB& b, m3d_vec3 corr)
{
	float theta = m3d_length(corr);
	if (theta < 1e-9f)
	{
		return;
	}
	m3d_vec3 n = m3d_scale(1.0f / theta, corr);
	float wA = a.invMass != 0.0f ? m3d_dot(n, m3d_mat3_mulv(&a.iw, n)) : 0.0f;
	float wB = b.invMass != 0.0f ? m3d_dot(n, m3d_mat3_mulv(&b.iw, n)) : 0.0f;
	float den = wA + wB;
	if (den <= 0.0f)
	{
		return;
	}
	float lambda = theta / den;
	m3d_vec3 p = m3d_scale(lambda, n);
	if (a.invMass != 0.0f)
	{
		a.q = rotate_by(a.q, m3d_mat3_mulv(&a.iw, p));
	}
	if (b.invMass != 0.0f)
	{
		m3d_vec3 dw = m3d_mat3_mulv(&b.iw, p);
		b.q = rotate_by(b.q, m3d_v3(-dw.x, -dw.y, -dw.z));
	}
}

// src/xpbd.cpp:477-543
MX_HD void solve_joint_position(DJoint& j, WB& a, WB& b, float h)
{
	if (j.type == 2)
	{
		m3d_vec3 axA = m3d_rotate(a.q, j.localAxisA);
		m3d_vec3 axB = m3d_rotate(b.q, j.localAxisB);
		apply_angular_correction(a, b, m3d_cross(axA, axB));
		if (j.enableLimit)
		{
			float phi = hinge_angle(a.q, b.q, j);
			float cl = phi < j.lowerAngle ? j.lowerAngle : (phi > j.upperAngle ? j.upperAngle : phi);
			if (cl != phi)
			{
				m3d_vec3 axis = m3d_rotate(a.q, j.localAxisA);
				apply_angular_correction(a, b, m3d_scale(phi - cl, axis));
			}
		}
	}
	else if (j.type == 3)
	{
		m3d_quat target = m3d_quat_mul(a.q, j.relRot0);
		m3d_quat qe = m3d_quat_mul(b.q, m3d_quat_conj(target));
		float sign = qe.w >= 0.0f ? 1.0f : -1.0f;
		apply_angular_correction(a, b, m3d_scale(2.0f * sign, m3d_v3(qe.x, qe.y, qe.z)));
	}
	m3d_vec3 rA = m3d_rotate(a.q, j.localAnchorA);
	m3d_vec3 rB = m3d_rotate(b.q, j.localAnchorB);
	m3d_vec3 d = m3d_sub(m3d_add(b.p, rB), m3d_add(a.p, rA));
	float len = m3d_length(d);
	if (len < 1e-9f)
	{
		return;
	}
	m3d_vec3 u = m3d_scale(1.0f / len, d);
	float Cv = j.type == 0 ? len - j.length : len;
	float w = generalized_inv_mass(a, rA, u) + generalized_inv_mass(b, rB, u);
	float alphaT = j.compliance / (h * h);
	float den = w + alphaT;
	if (den <= 0.0f)
	{
		return;
	}
	float dl = (-Cv - alphaT * j.lambda) / den;
	j.lambda += dl;
	m3d_vec3 p = m3d_scale(dl, u);
	apply_position_correction(b, p, rB);
	apply_position_correction(a, m3d_neg(p), rA);
}

// hinge motor; src/xpbd.cpp:549-581
MX_HD void solve_joint_velocity(DJoint& j, WB& a, WB& b, float h)
{
	if (j.type != 2 || !j.enableMotor || j.maxMotorTorque <= 0.0f)
	{
		return;
	}
	m3d_vec3 axis = m3d_rotate(a.q, j.localAxisA);
	float wA = a.invMass != 0.0f ? m3d_dot(axis, m3d_mat3_mulv(&a.iw, axis)) : 0.0f;
	float wB = b.invMass != 0.0f ? m3d_dot(axis, m3d_mat3_mulv(&b.iw, axis)) : 0.0f;
	float den = wA + wB;
	if (den <= 0.0f)
	{
		return;
	}
	float wrel = m3d_dot(m3d_sub(b.w, a.w), axis);
	float dl = (j.motorSpeed - wrel) / den;
	float budget = j.maxMotorTorque * h;
	float nl = j.motorLambda + dl;
	nl = nl > budget ? budget : (nl < -budget ? -budget : nl);
	dl = nl - j.motorLambda;
	j.motorLambda = nl;
	m3d_vec3 p = m3d_scale(dl, axis);
	if (b.invMass != 0.0f)
	{
		b.w = m3d_add(b.w, m3d_mat3_mulv(&b.iw, p));
	}
	if (a.invMass != 0.0f)
	{
		a.w = m3d_sub(a.w, m3d_mat3_mulv(&a.iw, p));
	}
}

} // namespace mx
