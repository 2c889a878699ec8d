/*
 * scene_driver.c - synthetic benchmark scenes driven through the m3d_* API,
 * in native code so that a streaming scene's per-step spawns and despawns
 * cost what they cost a C caller (the reference's bench does them in C++,
 * bench/bench.cpp:209-253).
 *
 * The driver calls the engine through a table of function pointers, so the
 * same code drives any library exporting the API (this backend, or the
 * reference build used as the CPU baseline).
 *
 * Random draws: the LCG of bench/bench.cpp:121-125 (s = s*1664525 +
 * 1013904223, u = (s >> 8) / 0xFFFFFF), in double precision; per-body draw
 * order as in melon3d_amd/scenes.py (position, size draws for the kind,
 * quaternion, friction), so a scene built here equals the Python builder's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "melon3d.h"

typedef struct m3ds_api
{
	m3d_body_def (*body_def_default)(void);
	m3d_shape_def (*shape_def_default)(void);
	m3d_body_id (*body_create)(m3d_world*, const m3d_body_def*, const m3d_shape_def*);
	void (*body_destroy)(m3d_world*, m3d_body_id);
	/* bulk despawn + spawn (m3dx_body_respawn), or NULL: one call per body */
	int (*respawn)(m3d_world*, const m3d_body_id*, const m3d_body_def*, const m3d_shape_def*, int32_t, m3d_body_id*);
} m3ds_api;

typedef struct m3ds_stream
{
	m3ds_api api;
	m3d_world* world;
	uint32_t rng;
	int32_t count;
	int32_t churn;
	int32_t cursor;
	float arena;
	m3d_body_id* ring;
	/* one churn step's arguments for a bulk respawn */
	m3d_body_id* gone;
	m3d_body_def* bds;
	m3d_shape_def* sds;
	m3d_body_id* born;
} m3ds_stream;

static double lcg_u(uint32_t* s)
{
	*s = *s * 1664525u + 1013904223u;
	return (double)(*s >> 8) / (double)0x00FFFFFF;
}

static double lcg_uniform(uint32_t* s, double a, double b) { return a + (b - a) * lcg_u(s); }

static const double kPi = 3.14159265358979323846;

static void mixed_defs(const m3ds_api* api, uint32_t* rng, int kind, double x, double y, double z, double vy,
					   m3d_body_def* bdOut, m3d_shape_def* sdOut)
{
	m3d_body_def bd = api->body_def_default();
	bd.type = M3D_BODY_DYNAMIC;
	bd.position.x = (float)x;
	bd.position.y = (float)y;
	bd.position.z = (float)z;
	bd.linearVelocity.x = 0.0f;
	bd.linearVelocity.y = (float)vy;
	bd.linearVelocity.z = 0.0f;
	m3d_shape_def sd = api->shape_def_default();
	if (kind == 0)
	{
		sd.type = M3D_SHAPE_SPHERE;
		sd.radius = (float)lcg_uniform(rng, 0.2, 0.5);
	}
	else if (kind == 1)
	{
		sd.type = M3D_SHAPE_CAPSULE;
		sd.radius = (float)lcg_uniform(rng, 0.15, 0.3);
		sd.halfHeight = (float)lcg_uniform(rng, 0.2, 0.6);
	}
	else
	{
		sd.type = M3D_SHAPE_BOX;
		double hx = lcg_uniform(rng, 0.2, 0.6);
		double hy = lcg_uniform(rng, 0.2, 0.6);
		double hz = lcg_uniform(rng, 0.2, 0.6);
		sd.halfExtents.x = (float)hx;
		sd.halfExtents.y = (float)hy;
		sd.halfExtents.z = (float)hz;
	}
	/* Shoemake's uniform rotation from three uniforms */
	double u1 = lcg_u(rng), u2 = lcg_u(rng), u3 = lcg_u(rng);
	double a = sqrt(1.0 - u1), b = sqrt(u1);
	bd.orientation.x = (float)(a * sin(2.0 * kPi * u2));
	bd.orientation.y = (float)(a * cos(2.0 * kPi * u2));
	bd.orientation.z = (float)(b * sin(2.0 * kPi * u3));
	bd.orientation.w = (float)(b * cos(2.0 * kPi * u3));
	sd.friction = (float)lcg_uniform(rng, 0.3, 0.8);
	*bdOut = bd;
	*sdOut = sd;
}

static m3d_body_id mixed_body(const m3ds_api* api, m3d_world* w, uint32_t* rng, int kind, double x, double y, double z,
							  double vy)
{
	m3d_body_def bd;
	m3d_shape_def sd;
	mixed_defs(api, rng, kind, x, y, z, vy, &bd, &sd);
	return api->body_create(w, &bd, &sd);
}

static void add_ground(const m3ds_api* api, m3d_world* w, float hx, float hy, float hz, float y)
{
	m3d_body_def bd = api->body_def_default();
	bd.type = M3D_BODY_STATIC;
	bd.position.x = 0.0f;
	bd.position.y = y;
	bd.position.z = 0.0f;
	m3d_shape_def sd = api->shape_def_default();
	sd.type = M3D_SHAPE_BOX;
	sd.halfExtents.x = hx;
	sd.halfExtents.y = hy;
	sd.halfExtents.z = hz;
	sd.friction = 0.6f;
	(void)api->body_create(w, &bd, &sd);
}

/* Rain pile: kind = i % 3, positions uniform in a gx * gy * gz column starting
 * 1 m above the ground (draws x, y, z, then the body's draws). */
int m3ds_rain(const m3ds_api* api, m3d_world* w, int count, uint32_t seed, float gx, float gz, float gy)
{
	add_ground(api, w, gx > 50.0f ? gx : 50.0f, 1.0f, gz > 50.0f ? gz : 50.0f, -1.0f);
	uint32_t rng = seed;
	for (int i = 0; i < count; ++i)
	{
		double x = lcg_uniform(&rng, -0.5 * gx, 0.5 * gx);
		double y = 1.0 + lcg_uniform(&rng, 0.0, gy);
		double z = lcg_uniform(&rng, -0.5 * gz, 0.5 * gz);
		(void)mixed_body(api, w, &rng, i % 3, x, y, z, 0.0);
	}
	return count;
}

/* Box pyramid rows of the reference bench (bench/bench.cpp:48-72): row r holds
 * base - r boxes of half-extent 0.5 centred on x = 0 at y = 0.5 + r; the
 * smallest base that holds `count` boxes, filled bottom-up and left to right
 * until `count` are placed. */
static void pyramid_rows(const m3ds_api* api, m3d_world* w, int count)
{
	int base = 1;
	while (base * (base + 1) / 2 < count)
		++base;
	int made = 0;
	for (int row = 0; row < base && made < count; ++row)
	{
		int n = base - row;
		double x0 = -0.5 * (double)(n - 1);
		for (int i = 0; i < n && made < count; ++i, ++made)
		{
			m3d_body_def bd = api->body_def_default();
			bd.type = M3D_BODY_DYNAMIC;
			bd.position.x = (float)(x0 + (double)i);
			bd.position.y = (float)(0.5 + (double)row);
			bd.position.z = 0.0f;
			m3d_shape_def sd = api->shape_def_default();
			sd.type = M3D_SHAPE_BOX;
			sd.halfExtents.x = 0.5f;
			sd.halfExtents.y = 0.5f;
			sd.halfExtents.z = 0.5f;
			sd.friction = 0.6f;
			(void)api->body_create(w, &bd, &sd);
		}
	}
}

/* One world of the batched workload, index w (seed 1000003 * (w + 1)):
 * dynamic body count floor(exp(U[ln 16, ln 2049))), then a pile (u < 1/2) or
 * a pyramid. The pile follows the rain rules in a column scaled to the rain
 * pile's density (40 x 60 x 40 m per 20 000 bodies, times (count/20000)^(1/3))
 * on the rain ground; the pyramid stands on the pyramid ground (half-extents
 * 50 x 0.5 x 50, top face at y = 0). Returns the dynamic body count; *isPile
 * receives the kind. */
int m3ds_batch_world(const m3ds_api* api, m3d_world* w, int index, int* isPile)
{
	uint32_t rng = 1000003u * (uint32_t)(index + 1);
	double lo = log(16.0), hi = log(2049.0);
	int count = (int)floor(exp(lcg_uniform(&rng, lo, hi)));
	if (count < 16)
		count = 16;
	if (count > 2048)
		count = 2048;
	int pile = lcg_u(&rng) < 0.5;
	if (isPile)
		*isPile = pile;
	if (pile)
	{
		double sc = cbrt((double)count / 20000.0);
		double gx = 40.0 * sc, gz = 40.0 * sc, gy = 60.0 * sc;
		add_ground(api, w, gx > 50.0 ? (float)gx : 50.0f, 1.0f, gz > 50.0 ? (float)gz : 50.0f, -1.0f);
		for (int i = 0; i < count; ++i)
		{
			double x = lcg_uniform(&rng, -0.5 * gx, 0.5 * gx);
			double y = 1.0 + lcg_uniform(&rng, 0.0, gy);
			double z = lcg_uniform(&rng, -0.5 * gz, 0.5 * gz);
			(void)mixed_body(api, w, &rng, i % 3, x, y, z, 0.0);
		}
	}
	else
	{
		add_ground(api, w, 50.0f, 0.5f, 50.0f, -0.5f);
		pyramid_rows(api, w, count);
	}
	return count;
}

void m3ds_stream_free(m3ds_stream* st);

static int alloc_churn(m3ds_stream* st, int churn)
{
	size_t n = (size_t)(churn > 0 ? churn : 1);
	st->gone = (m3d_body_id*)malloc(sizeof(m3d_body_id) * n);
	st->bds = (m3d_body_def*)malloc(sizeof(m3d_body_def) * n);
	st->sds = (m3d_shape_def*)malloc(sizeof(m3d_shape_def) * n);
	st->born = (m3d_body_id*)malloc(sizeof(m3d_body_id) * n);
	return st->gone && st->bds && st->sds && st->born;
}

/* Streaming scene: count bodies uniform in an arena x 50 x arena volume; each
 * churn step destroys the churn oldest (FIFO ring) and spawns as many at
 * y in [30, 60] with downward speed U[5, 20] m/s. */
m3ds_stream* m3ds_stream_create(const m3ds_api* api, m3d_world* w, int count, int churn, uint32_t seed, float arena)
{
	m3ds_stream* st = (m3ds_stream*)calloc(1, sizeof(m3ds_stream));
	if (!st)
		return NULL;
	st->ring = (m3d_body_id*)malloc(sizeof(m3d_body_id) * (size_t)(count > 0 ? count : 1));
	if (!st->ring || !alloc_churn(st, churn))
	{
		m3ds_stream_free(st);
		return NULL;
	}
	st->api = *api;
	st->world = w;
	st->rng = seed;
	st->count = count;
	st->churn = churn;
	st->arena = arena;
	add_ground(api, w, 0.5f * arena + 10.0f, 1.0f, 0.5f * arena + 10.0f, -1.0f);
	for (int i = 0; i < count; ++i)
	{
		double x = lcg_uniform(&st->rng, -0.5 * arena, 0.5 * arena);
		double y = 1.0 + lcg_uniform(&st->rng, 0.0, 50.0);
		double z = lcg_uniform(&st->rng, -0.5 * arena, 0.5 * arena);
		st->ring[i] = mixed_body(api, w, &st->rng, i % 3, x, y, z, 0.0);
	}
	return st;
}

void m3ds_stream_churn(m3ds_stream* st)
{
	if (st->api.respawn != NULL && st->count > 0)
	{
		/* the same draws in the same order, handed over in one call */
		int n = st->churn;
		int cur = st->cursor;
		for (int k = 0; k < n; ++k)
		{
			st->gone[k] = st->ring[cur];
			double x = lcg_uniform(&st->rng, -0.5 * st->arena, 0.5 * st->arena);
			double y = lcg_uniform(&st->rng, 30.0, 60.0);
			double z = lcg_uniform(&st->rng, -0.5 * st->arena, 0.5 * st->arena);
			double vy = -lcg_uniform(&st->rng, 5.0, 20.0);
			mixed_defs(&st->api, &st->rng, cur % 3, x, y, z, vy, &st->bds[k], &st->sds[k]);
			cur = (cur + 1) % st->count;
		}
		st->api.respawn(st->world, st->gone, st->bds, st->sds, n, st->born);
		for (int k = 0; k < n; ++k)
		{
			st->ring[st->cursor] = st->born[k];
			st->cursor = (st->cursor + 1) % st->count;
		}
		return;
	}
	for (int k = 0; k < st->churn && st->count > 0; ++k)
	{
		st->api.body_destroy(st->world, st->ring[st->cursor]);
		double x = lcg_uniform(&st->rng, -0.5 * st->arena, 0.5 * st->arena);
		double y = lcg_uniform(&st->rng, 30.0, 60.0);
		double z = lcg_uniform(&st->rng, -0.5 * st->arena, 0.5 * st->arena);
		double vy = -lcg_uniform(&st->rng, 5.0, 20.0);
		st->ring[st->cursor] = mixed_body(&st->api, st->world, &st->rng, st->cursor % 3, x, y, z, vy);
		st->cursor = (st->cursor + 1) % st->count;
	}
}

/* The same stream state continued on another world holding the same bodies
 * (a world imported from this one's snapshot), through another library. */
m3ds_stream* m3ds_stream_clone(const m3ds_stream* src, const m3ds_api* api, m3d_world* w)
{
	m3ds_stream* st = (m3ds_stream*)calloc(1, sizeof(m3ds_stream));
	if (!st)
		return NULL;
	st->ring = (m3d_body_id*)malloc(sizeof(m3d_body_id) * (size_t)(src->count > 0 ? src->count : 1));
	if (!st->ring || !alloc_churn(st, src->churn))
	{
		m3ds_stream_free(st);
		return NULL;
	}
	memcpy(st->ring, src->ring, sizeof(m3d_body_id) * (size_t)src->count);
	st->api = *api;
	st->world = w;
	st->rng = src->rng;
	st->count = src->count;
	st->churn = src->churn;
	st->cursor = src->cursor;
	st->arena = src->arena;
	return st;
}

void m3ds_stream_free(m3ds_stream* st)
{
	if (st)
	{
		free(st->ring);
		free(st->gone);
		free(st->bds);
		free(st->sds);
		free(st->born);
		free(st);
	}
}
