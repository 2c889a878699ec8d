/*
 * m3d_oracle.c - CPU restatement of the reference's discrete step stages.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ (and nothing else) as the
 * checker for the device backend's integer/index outputs. Never linked into
 * or called by libmelon3d_amd.so.
 *
 * Each function names the reference code it restates. Pins (see tests/):
 *   orc_cell_key        closed form + reference grid dumps      grid.cpp:19-37
 *   orc_build_tier      reference sorted tiers, bit-exact       grid.cpp:39-131
 *   orc_brute_pairs     reference created-pair sets             world.cpp:1486-1702
 *   orc_islands         reference island labels; BFS property   world.cpp:1868-1958
 *   orc_greedy_color    reference color buckets                 xpbd.cpp:1179-1235
 *   orc_color_valid     definition (no dynamic body twice)      xpbd.cpp:1179-1181
 *   orc_levels          definition (serial-order equivalence)   xpbd.cpp:923-973
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_entry
{
	uint64_t key;
	int32_t body;
	int32_t pad;
} orc_entry;

/* grid.cpp:19-31: floor(v * (1/cell)) as int64, +2^20, 21 bits per axis */
static int64_t cell_coord(float v, float inv) { return (int64_t)floorf(v * inv); }

static uint64_t pack_cell(int64_t x, int64_t y, int64_t z)
{
	const int64_t off = (int64_t)1 << 20;
	uint64_t ux = (uint64_t)(x + off) & 0x1FFFFFu;
	uint64_t uy = (uint64_t)(y + off) & 0x1FFFFFu;
	uint64_t uz = (uint64_t)(z + off) & 0x1FFFFFu;
	return (ux << 42) | (uy << 21) | uz;
}

uint64_t orc_cell_key(float x, float y, float z, float cellSize)
{
	float inv = 1.0f / cellSize;
	return pack_cell(cell_coord(x, inv), cell_coord(y, inv), cell_coord(z, inv));
}

static int entry_cmp(const void* pa, const void* pb)
{
	const orc_entry* a = (const orc_entry*)pa;
	const orc_entry* b = (const orc_entry*)pb;
	if (a->key != b->key)
	{
		return a->key < b->key ? -1 : 1;
	}
	return (a->body > b->body) - (a->body < b->body);
}

/* grid.cpp:39-95. aabbs: 6 floats per body (lower xyz, upper xyz).
 * Returns the entry count (entries beyond cap are counted, not written);
 * *oversizedCount gets the coarse-list length; inGrid[i] = 1 for bodies with
 * entries. Entries are sorted by (key, body). */
int64_t orc_build_tier(int32_t n, const float* aabbs, const uint8_t* include, float cellSize, orc_entry* entries,
					   int64_t cap, int32_t* oversized, int32_t* oversizedCount, uint8_t* inGrid)
{
	float inv = 1.0f / cellSize;
	int64_t count = 0;
	int32_t ov = 0;
	for (int32_t i = 0; i < n; ++i)
	{
		inGrid[i] = 0;
		if (!include[i])
		{
			continue;
		}
		const float* b = aabbs + 6 * (int64_t)i;
		int64_t x0 = cell_coord(b[0], inv), y0 = cell_coord(b[1], inv), z0 = cell_coord(b[2], inv);
		int64_t x1 = cell_coord(b[3], inv), y1 = cell_coord(b[4], inv), z1 = cell_coord(b[5], inv);
		int64_t span = (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
		if (span > 64)
		{
			oversized[ov++] = i;
			continue;
		}
		inGrid[i] = 1;
		for (int64_t x = x0; x <= x1; ++x)
			for (int64_t y = y0; y <= y1; ++y)
				for (int64_t z = z0; z <= z1; ++z)
				{
					if (count < cap)
					{
						entries[count].key = pack_cell(x, y, z);
						entries[count].body = i;
						entries[count].pad = 0;
					}
					++count;
				}
	}
	if (count <= cap)
	{
		qsort(entries, (size_t)count, sizeof(orc_entry), entry_cmp);
	}
	*oversizedCount = ov;
	return count;
}

static int aabb_overlap(const float* a, const float* b)
{
	return a[0] <= b[3] && a[1] <= b[4] && a[2] <= b[5] && b[0] <= a[3] && b[1] <= a[4] && b[2] <= a[5];
}

/* Per-body admission inputs for orc_brute_pairs (world.cpp:1486-1538). */
typedef struct orc_body_flags
{
	uint8_t inUse;
	uint8_t dynamic;
	uint8_t awake;
	uint8_t moved;
	uint8_t activeTier;  /* in the active tier (grid cells or active coarse list) */
	uint8_t stableTier;  /* in the stable tier (cells or coarse list) */
	uint8_t pad0, pad1;
	uint32_t category;
	uint32_t mask;
	int32_t group;
} orc_body_flags;

static int key_cmp(const void* pa, const void* pb)
{
	uint64_t a = *(const uint64_t*)pa, b = *(const uint64_t*)pb;
	return (a > b) - (a < b);
}

static int key_in(const uint64_t* sorted, int64_t n, uint64_t k)
{
	return n > 0 && bsearch(&k, sorted, (size_t)n, sizeof(uint64_t), key_cmp) != NULL;
}

/* Brute-force restatement of the pair set the step creates: every pair
 * with at least one active-tier member whose fat AABBs overlap and that
 * passes the admission filter (world.cpp:1486-1538), excluding existing
 * contacts and jointed pairs (both sorted key lists). Writes ascending
 * lo<<32|hi keys; returns the count. This is the created SET; creation
 * ORDER is pinned separately against the reference build. */
int64_t orc_brute_pairs(int32_t n, const float* aabbs, const orc_body_flags* f, const uint64_t* existing,
						int64_t existingCount, const uint64_t* jointed, int64_t jointedCount, uint64_t* out,
						int64_t cap)
{
	int64_t count = 0;
	for (int32_t a = 0; a < n; ++a)
	{
		for (int32_t b = a + 1; b < n; ++b)
		{
			const orc_body_flags* fa = f + a;
			const orc_body_flags* fb = f + b;
			if (!(fa->activeTier || fb->activeTier))
				continue;
			if (!((fa->activeTier || fa->stableTier) && (fb->activeTier || fb->stableTier)))
				continue;
			if (!fa->moved && !fb->moved)
				continue;
			if (!fa->inUse || !fb->inUse)
				continue;
			if (!fa->dynamic && !fb->dynamic)
				continue;
			if (!fa->awake && !fb->awake)
				continue;
			if (!aabb_overlap(aabbs + 6 * (int64_t)a, aabbs + 6 * (int64_t)b))
				continue;
			if (fa->group != 0 && fa->group == fb->group)
			{
				if (fa->group < 0)
					continue;
			}
			else if ((fa->category & fb->mask) == 0 || (fb->category & fa->mask) == 0)
			{
				continue;
			}
			uint64_t k = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
			if (key_in(existing, existingCount, k) || key_in(jointed, jointedCount, k))
				continue;
			if (count < cap)
				out[count] = k;
			++count;
		}
	}
	return count;
}

/* world.cpp:1868-1958 restated as BFS connected components over touching,
 * non-sensor, dynamic-dynamic contacts and dynamic-dynamic joints. Islands
 * are numbered in order of their smallest body index (the reference assigns
 * island slots while scanning bodies ascending). islandOf[i] = -1 for
 * non-dynamic bodies. Returns the island count. */
int32_t orc_islands(int32_t n, const uint8_t* dynamic, int32_t edgeCount, const int32_t* edgeA,
					const int32_t* edgeB, int32_t* islandOf)
{
	int32_t* head = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
	int32_t* next = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * edgeCount + 1));
	int32_t* to = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * edgeCount + 1));
	int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
	for (int32_t i = 0; i < n; ++i)
	{
		head[i] = -1;
		islandOf[i] = -1;
	}
	int32_t m = 0;
	for (int32_t e = 0; e < edgeCount; ++e)
	{
		int32_t a = edgeA[e], b = edgeB[e];
		if (!dynamic[a] || !dynamic[b])
			continue;
		to[m] = b;
		next[m] = head[a];
		head[a] = m++;
		to[m] = a;
		next[m] = head[b];
		head[b] = m++;
	}
	int32_t islands = 0;
	for (int32_t s = 0; s < n; ++s)
	{
		if (!dynamic[s] || islandOf[s] >= 0)
			continue;
		int32_t qh = 0, qt = 0;
		queue[qt++] = s;
		islandOf[s] = islands;
		while (qh < qt)
		{
			int32_t x = queue[qh++];
			for (int32_t k = head[x]; k >= 0; k = next[k])
			{
				int32_t y = to[k];
				if (islandOf[y] < 0)
				{
					islandOf[y] = islands;
					queue[qt++] = y;
				}
			}
		}
		++islands;
	}
	free(head);
	free(next);
	free(to);
	free(queue);
	return islands;
}

/* xpbd.cpp:1179-1235: sequential first-fit over contacts in island order.
 * dynA/dynB: endpoint has invMass > 0. Colors 0..7; 8 = overflow.
 * colorMask is caller scratch of bodyCount words. */
void orc_greedy_color(int32_t count, const int32_t* bodyA, const int32_t* bodyB, const uint8_t* dynA,
					  const uint8_t* dynB, int32_t bodyCount, uint32_t* colorMask, int32_t* color)
{
	memset(colorMask, 0, sizeof(uint32_t) * (size_t)bodyCount);
	for (int32_t k = 0; k < count; ++k)
	{
		uint32_t used = 0;
		if (dynA[k])
			used |= colorMask[bodyA[k]];
		if (dynB[k])
			used |= colorMask[bodyB[k]];
		uint32_t freeBits = ~used;
		int32_t c = 0;
		while (c < 32 && !((freeBits >> c) & 1u))
			++c;
		if (c >= 8)
		{
			color[k] = 8;
			continue;
		}
		color[k] = c;
		if (dynA[k])
			colorMask[bodyA[k]] |= 1u << c;
		if (dynB[k])
			colorMask[bodyB[k]] |= 1u << c;
	}
}

/* Validity: within each color 0..7 no dynamic body appears twice.
 * Returns 1 when valid. */
int32_t orc_color_valid(int32_t count, const int32_t* bodyA, const int32_t* bodyB, const uint8_t* dynA,
						const uint8_t* dynB, const int32_t* color, int32_t bodyCount)
{
	uint8_t* seen = (uint8_t*)calloc((size_t)bodyCount * 8, 1);
	int32_t ok = 1;
	for (int32_t k = 0; k < count && ok; ++k)
	{
		int32_t c = color[k];
		if (c < 0 || c >= 8)
			continue;
		if (dynA[k])
		{
			if (seen[(size_t)bodyA[k] * 8 + c])
				ok = 0;
			seen[(size_t)bodyA[k] * 8 + c] = 1;
		}
		if (dynB[k])
		{
			if (seen[(size_t)bodyB[k] * 8 + c])
				ok = 0;
			seen[(size_t)bodyB[k] * 8 + c] = 1;
		}
	}
	free(seen);
	return ok;
}

/* Level of each constraint in a serial Gauss-Seidel order: one more than the
 * deepest earlier constraint sharing a written (dynamic) body. Solving level
 * by level reproduces the serial order's per-body update sequence exactly
 * (xpbd.cpp:923-973 loops contacts in order). */
void orc_levels(int32_t count, const int32_t* bodyA, const int32_t* bodyB, const uint8_t* dynA, const uint8_t* dynB,
				int32_t bodyCount, int32_t* lastLevel, int32_t* level)
{
	for (int32_t i = 0; i < bodyCount; ++i)
		lastLevel[i] = -1;
	for (int32_t k = 0; k < count; ++k)
	{
		int32_t l = -1;
		if (dynA[k] && lastLevel[bodyA[k]] > l)
			l = lastLevel[bodyA[k]];
		if (dynB[k] && lastLevel[bodyB[k]] > l)
			l = lastLevel[bodyB[k]];
		level[k] = l + 1;
		if (dynA[k])
			lastLevel[bodyA[k]] = l + 1;
		if (dynB[k])
			lastLevel[bodyB[k]] = l + 1;
	}
}
