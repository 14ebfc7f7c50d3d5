/*
 * oracle.c -- CPU restatement of Nebula's electron-tracking path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compile with -ffp-contract=off and
 * without -ffast-math: every floating-point expression below keeps the
 * operation order of the reference function cited beside it, so results are
 * bit-identical to the reference's CPU build.  Reference paths are relative
 * to its source/ directory.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* config/data_types.h:55 -- absolute tolerance in nm. */
#define EPS (10.0f * FLT_EPSILON)
/* config/config.h:32 */
static const float PI_F = 3.1415926535897932f;

/* Cap on octree cells visited by one propagate call and on plasmon-decay
 * rejection rounds.  The reference loops are unbounded; the HIP kernels need
 * a bound every wave reaches, and the oracle applies the same bound. */
#define ORC_MAX_CELLS 1000000u
#define ORC_MAX_REJECT 1000

/* ------------------------------------------------------------- math mode -- */
/* Transcendentals of the tracking loop.  Material preprocessing (create)
 * always uses the C library, as the host code of the reference does.
 *
 * ORC_MATH_PF: the float transcendentals of the HIP kernels, restated.  They
 * use only IEEE float +, -, *, / and integer bit operations, so with
 * -ffp-contract=off the CPU and the GPU produce the same bits.  Methods:
 * exponent/mantissa split with the log(1+f) = f - f^2/2 + s(f^2/2 + R(s^2))
 * identity (s = f/(2+f), R the atanh series); Cody-Waite reduction by
 * k*ln2 or k*pi/2 with split constants and Taylor polynomials; arctangent
 * by reduction to |u| <= tan(pi/8) and a least-squares polynomial; cube root
 * by a bit-level first guess and four Newton steps.  Accuracy against the
 * exact functions is checked by tests/test_math.py. */
#ifdef ORC_MATH_PF
#define PF_P1 0x1.922p+0f
#define PF_P2 -0x1.2aep-18f
#define PF_P3 -0x1.de973ep-31f
#define PF_LN2_HI 0x1.62e4p-1f
#define PF_LN2_LO 0x1.7f7d1cp-20f
#define PF_INV_LN2 0x1.715476p+0f
#define PF_PI_HI 0x1.921fb6p+1f
#define PF_PI_LO -0x1.777a5cp-24f
#define PF_PI_2_HI 0x1.921fb6p+0f
#define PF_PI_2_LO -0x1.777a5cp-25f
#define PF_PI_4_HI 0x1.921fb6p-1f
#define PF_PI_4_LO -0x1.777a5cp-26f
#define PF_TWO_OVER_PI 0x1.45f306p-1f
#define PF_TAN_PI_8 0x1.a8279ap-2f

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline float pf_nan(void) { return u2f(0x7fc00000u); }
/* 2^k for -126 <= k <= 127 */
static inline float pf_pow2(int k) { return u2f((uint32_t)(k + 127) << 23); }

static float pf_log(float x)
{
	if (x != x) return x;
	if (x < 0.0f) return pf_nan();
	if (x == 0.0f) return -INFINITY;
	uint32_t i = f2u(x);
	if (i >= 0x7f800000u) return x;
	int e = 0;
	if (i < 0x00800000u) { x = x * 0x1p25f; i = f2u(x); e = -25; }
	e += (int)(i >> 23) - 127;
	const uint32_t m = i & 0x007fffffu;
	if (m > 0x3504f3u) { i = m | 0x3f000000u; e += 1; }   /* mantissa above sqrt(2): halve */
	else i = m | 0x3f800000u;
	const float f = u2f(i) - 1.0f;
	const float s = f / (2.0f + f);
	const float z = s * s;
	const float R = z * (0x1.555556p-1f + z * (0x1.99999ap-2f + z * (0x1.24924ap-2f + z * 0x1.c71c72p-3f)));
	const float hfsq = 0.5f * f * f;
	const float l1p = f - (hfsq - s * (hfsq + R));
	const float fe = (float)e;
	return fe * PF_LN2_HI + (l1p + fe * PF_LN2_LO);
}

/* (exp(r) - 1 - r) / r^2, |r| <= ln2/2 */
static inline float pf_exp_p(float r)
{
	return 0.5f + r * (0x1.555556p-3f + r * (0x1.555556p-5f + r * (0x1.111112p-7f
		+ r * (0x1.6c16c2p-10f + r * 0x1.a01a02p-13f))));
}

static float pf_exp(float x)
{
	if (x != x) return x;
	if (x > 0x1.62e42ep+6f) return INFINITY;
	if (x < -0x1.9fe368p+6f) return 0.0f;
	const float kf = rintf(x * PF_INV_LN2);
	const int k = (int)kf;
	const float r = (x - kf * PF_LN2_HI) - kf * PF_LN2_LO;
	const float e = 1.0f + (r + (r * r) * pf_exp_p(r));
	if (k > 127) return (e * 0x1p127f) * pf_pow2(k - 127);
	if (k < -126) return (e * 0x1p-126f) * pf_pow2(k + 126);
	return e * pf_pow2(k);
}

static float pf_expm1(float x)
{
	if (x != x) return x;
	if (x > 0x1.62e42ep+6f) return INFINITY;
	if (x < -0x1.154246p+4f) return -1.0f;     /* below -25 ln2 */
	if (fabsf(x) < 0x1.62e430p-2f) {           /* |x| < ln2/2 */
		if (fabsf(x) < 0x1p-28f) return x;
		return x + (x * x) * pf_exp_p(x);
	}
	const float kf = rintf(x * PF_INV_LN2);
	const int k = (int)kf;
	const float r = (x - kf * PF_LN2_HI) - kf * PF_LN2_LO;
	const float em = r + (r * r) * pf_exp_p(r);
	if (k > 127) return ((em + 1.0f) * 0x1p127f) * 2.0f - 1.0f;
	const float t = pf_pow2(k);
	return em * t + (t - 1.0f);
}

static float pf_cbrt(float x)
{
	if (x != x || x == 0.0f || fabsf(x) == INFINITY) return x;
	float ax = fabsf(x);
	float post = 1.0f;
	if (ax < 0x1p-126f) { ax = ax * 0x1p24f; post = 0x1p-8f; }
	float y = u2f(f2u(ax) / 3u + 0x2a555555u);
	for (int it = 0; it < 4; ++it)
		y = (2.0f * y + ax / (y * y)) / 3.0f;
	y = y * post;
	return copysignf(y, x);
}

/* sin and cos of the reduced argument, |r| <= pi/4 */
static inline float pf_sin_r(float r)
{
	const float z = r * r;
	return r + (r * z) * (-0x1.555556p-3f + z * (0x1.111112p-7f + z * (-0x1.a01a02p-13f
		+ z * (0x1.71de3ap-19f + z * -0x1.ae6456p-26f))));
}
static inline float pf_cos_r(float r)
{
	const float z = r * r;
	return 1.0f - (0.5f * z - (z * z) * (0x1.555556p-5f + z * (-0x1.6c16c2p-10f
		+ z * (0x1.a01a02p-16f + z * -0x1.27e4fcp-22f))));
}
/* x = k pi/2 + r */
static inline float pf_reduce(float x, int* q)
{
	const float kf = rintf(x * PF_TWO_OVER_PI);
	*q = ((int)kf) & 3;
	return ((x - kf * PF_P1) - kf * PF_P2) - kf * PF_P3;
}
static float pf_sin(float x)
{
	if (x != x || fabsf(x) == INFINITY) return pf_nan();
	int q;
	const float r = pf_reduce(x, &q);
	switch (q) {
	case 0: return pf_sin_r(r);
	case 1: return pf_cos_r(r);
	case 2: return -pf_sin_r(r);
	default: return -pf_cos_r(r);
	}
}
static float pf_cos(float x)
{
	if (x != x || fabsf(x) == INFINITY) return pf_nan();
	int q;
	const float r = pf_reduce(x, &q);
	switch (q) {
	case 0: return pf_cos_r(r);
	case 1: return -pf_sin_r(r);
	case 2: return -pf_cos_r(r);
	default: return pf_sin_r(r);
	}
}

/* atan(u), |u| <= tan(pi/8) */
static inline float pf_atan_r(float u)
{
	const float z = u * u;
	return u + (u * z) * (-0x1.555556p-2f + z * (0x1.999982p-3f + z * (-0x1.248b8p-3f
		+ z * (0x1.c5adfep-4f + z * (-0x1.62ae6ep-4f + z * 0x1.a49afap-5f)))));
}
static float pf_atan2(float y, float x)
{
	if (x != x || y != y) return x + y;
	const float ax = fabsf(x), ay = fabsf(y);
	const int xneg = signbit(x) != 0;
	if (ay == 0.0f)
		return xneg ? copysignf(PF_PI_HI, y) : y;
	if (ax == 0.0f)
		return copysignf(PF_PI_2_HI, y);
	if (ax == INFINITY || ay == INFINITY) {
		float a;
		if (ax == INFINITY && ay == INFINITY) a = xneg ? 3.0f * PF_PI_4_HI : PF_PI_4_HI;
		else if (ax == INFINITY) a = xneg ? PF_PI_HI : 0.0f;
		else a = PF_PI_2_HI;
		return copysignf(a, y);
	}
	const int swap = ay > ax;
	const float t = swap ? ax / ay : ay / ax;
	float a;
	if (t > PF_TAN_PI_8)
		a = PF_PI_4_HI + (pf_atan_r((t - 1.0f) / (t + 1.0f)) + PF_PI_4_LO);
	else
		a = pf_atan_r(t);
	if (swap) a = (PF_PI_2_HI - a) + PF_PI_2_LO;
	if (xneg) a = (PF_PI_HI - a) + PF_PI_LO;
	return copysignf(a, y);
}

static inline float m_log(float x) { return pf_log(x); }
static inline float m_exp(float x) { return pf_exp(x); }
static inline float m_expm1(float x) { return pf_expm1(x); }
static inline float m_cbrt(float x) { return pf_cbrt(x); }
static inline float m_sin(float x) { return pf_sin(x); }
static inline float m_cos(float x) { return pf_cos(x); }
static inline float m_atan2(float y, float x) { return pf_atan2(y, x); }
int orc_math_mode(void) { return 2; }

float orc_pf_eval(int fn, float a, float b)
{
	switch (fn) {
	case 0: return pf_log(a);
	case 1: return pf_exp(a);
	case 2: return pf_expm1(a);
	case 3: return pf_cbrt(a);
	case 4: return pf_sin(a);
	case 5: return pf_cos(a);
	default: return pf_atan2(a, b);
	}
}
#else
static inline float m_log(float x) { return logf(x); }
static inline float m_exp(float x) { return expf(x); }
static inline float m_expm1(float x) { return expm1f(x); }
static inline float m_cbrt(float x) { return cbrtf(x); }
static inline float m_sin(float x) { return sinf(x); }
static inline float m_cos(float x) { return cosf(x); }
static inline float m_atan2(float y, float x) { return atan2f(y, x); }
int orc_math_mode(void) { return 0; }

float orc_pf_eval(int fn, float a, float b)
{
	switch (fn) {
	case 0: return logf(a);
	case 1: return expf(a);
	case 2: return expm1f(a);
	case 3: return cbrtf(a);
	case 4: return sinf(a);
	case 5: return cosf(a);
	default: return atan2f(a, b);
	}
}
#endif

/* common/scalar_math.h:16-25 */
static inline int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
static inline float clampr(float i, float lo, float hi) { return i < lo ? lo : (i > hi ? hi : i); }
/* std::min / std::max semantics (CPU branch of minr/maxr). */
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }
static inline int floorr(float x) { return (int)floorf(x); }
static inline int rintr(float x) { return (int)rintf(x); }
static inline float saturater(float x) { return clampr(x, 0.0f, 1.0f); }

/* ---------------------------------------------------------- vector math -- */
/* common/vector_math.h */
static inline orc_vec3 v_add(orc_vec3 a, orc_vec3 b) { orc_vec3 r = { a.x + b.x, a.y + b.y, a.z + b.z }; return r; }
static inline orc_vec3 v_sub(orc_vec3 a, orc_vec3 b) { orc_vec3 r = { a.x - b.x, a.y - b.y, a.z - b.z }; return r; }
static inline orc_vec3 v_mul(orc_vec3 a, float b) { orc_vec3 r = { a.x * b, a.y * b, a.z * b }; return r; }
static inline orc_vec3 s_mul(float a, orc_vec3 b) { orc_vec3 r = { a * b.x, a * b.y, a * b.z }; return r; }
static inline orc_vec3 cross(orc_vec3 a, orc_vec3 b)
{
	orc_vec3 r = { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
	return r;
}
static inline float dot(orc_vec3 a, orc_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float mag2(orc_vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
/* normalise / normalised, CPU branch (vector_math.h:124-159): a * (1/sqrt). */
static inline orc_vec3 normalised(orc_vec3 a)
{
	const float r = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
	return v_mul(a, r);
}
/* vector_math.h:171-186.  Under the portable-float policy (the HIP
 * kernels' arithmetic) the cosine and sine of the azimuth atan2(y, x) are
 * taken as x/r and y/r with r = sqrt(x^2 + y^2), equal to the
 * transcendental form up to rounding; r = 0 takes the atan2 limits (cosine
 * +-1 by the sign of x, sine 0). */
static orc_vec3 make_normal_vec(orc_vec3 dir, float phi)
{
#ifdef ORC_MATH_PF
	const float r = sqrtf(dir.x * dir.x + dir.y * dir.y);
	float sin_az = 0.0f, cos_az = signbit(dir.x) ? -1.0f : 1.0f;
	if (r > 0.0f) {
		cos_az = dir.x / r;
		sin_az = dir.y / r;
	}
#else
	const float az = m_atan2(dir.y, dir.x);
	const float sin_az = m_sin(az), cos_az = m_cos(az);
#endif
	const orc_vec3 unit_v = { dir.z * cos_az, dir.z * sin_az, -sqrtf(dir.x * dir.x + dir.y * dir.y) };
	const orc_vec3 unit_u = cross(unit_v, dir);
	const float sin_phi = m_sin(phi), cos_phi = m_cos(phi);
	return v_add(v_mul(unit_u, cos_phi), v_mul(unit_v, sin_phi));
}
/* vector_math.h:194-206 */
static orc_vec3 make_unit_vec(float cos_theta, float phi)
{
	const float sin_phi = m_sin(phi), cos_phi = m_cos(phi);
	cos_theta = clampr(cos_theta, -1.0f, 1.0f);
	const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
	orc_vec3 r = { sin_theta * cos_phi, sin_theta * sin_phi, cos_theta };
	return r;
}

/* ------------------------------------------------------------------ RNG -- */
/* std::mt19937 (C++ [rand.eng.mers], parameters of mt19937). */
void orc_mt_seed(orc_rng* r, uint32_t seed)
{
	r->mt[0] = seed;
	for (int i = 1; i < 624; ++i)
		r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
	r->mti = 624;
}

uint32_t orc_mt_next(orc_rng* r)
{
	if (r->mti >= 624) {
		for (int i = 0; i < 624; ++i) {
			const uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
			r->mt[i] = r->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
		}
		r->mti = 0;
	}
	uint32_t z = r->mt[r->mti++];
	z ^= z >> 11;
	z ^= (z << 7) & 0x9d2c5680u;
	z ^= (z << 15) & 0xefc60000u;
	z ^= z >> 18;
	return z;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11; Random123 constants). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for (int round = 0; round < 10; ++round) {
		const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
		const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
		const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
		const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
		c0 = hi1 ^ c1 ^ k0;
		c1 = lo1;
		c2 = hi0 ^ c3 ^ k1;
		c3 = lo0;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_rng_init_mt(orc_rng* r, uint32_t seed)
{
	memset(r, 0, sizeof(*r));
	r->kind = 0;
	orc_mt_seed(r, seed);
}

void orc_rng_init_philox(orc_rng* r, uint64_t seed)
{
	memset(r, 0, sizeof(*r));
	r->kind = 1;
	r->key[0] = (uint32_t)seed;
	r->key[1] = (uint32_t)(seed >> 32);
}

void orc_rng_set_stream(orc_rng* r, uint64_t sid, uint32_t step)
{
	r->ctr[0] = (uint32_t)sid;
	r->ctr[1] = (uint32_t)(sid >> 32);
	r->ctr[2] = step;
	r->ctr[3] = 0;
	r->idx = 0;
}

void orc_rng_init_array(orc_rng* r, const float* u, size_t n)
{
	memset(r, 0, sizeof(*r));
	r->kind = 2;
	r->arr = u;
	r->arr_n = n;
}

/* Uniform draw.  kind 0 restates std::generate_canonical<float, 24>
 * (libstdc++ random.tcc) over mt19937: one 32-bit word, converted to float,
 * divided by 2^32, and pulled below 1 if it rounded up to 1.  kind 1 maps the
 * top 24 bits of a Philox word to (0, 1] exactly: ((w >> 8) + 1) * 2^-24. */
float orc_rng_unit(orc_rng* r)
{
	r->draws++;
	if (r->kind == 0) {
		const float sum = (float)orc_mt_next(r);
		float ret = sum / 4294967296.0f;
		if (ret >= 1.0f)
			ret = nextafterf(1.0f, 0.0f);
		return ret;
	}
	if (r->kind == 1) {
		if ((r->idx & 3u) == 0) {
			uint32_t c[4] = { r->ctr[0], r->ctr[1], r->ctr[2], r->idx >> 2 };
			orc_philox4x32_10(c, r->key, r->buf);
		}
		const uint32_t w = r->buf[r->idx & 3u];
		r->idx++;
		return (float)((w >> 8) + 1u) * (1.0f / 16777216.0f);
	}
	if (r->arr_pos < r->arr_n)
		return r->arr[r->arr_pos++];
	return 0.5f;
}

/* Stream id of a secondary electron: a splitmix64 finaliser of the parent's
 * (sid, step); the top bit marks secondaries so they never collide with the
 * primaries' sids (which are primary indices). */
uint64_t orc_se_sid(uint64_t parent_sid, uint32_t parent_step)
{
	uint64_t z = parent_sid * 0x9E3779B97F4A7C15ull + (uint64_t)parent_step + 1u;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	z ^= z >> 31;
	return z | 0x8000000000000000ull;
}

/* random.h:56-84 (uniform_vector draws phi first: g++ evaluates the
 * arguments of make_unit_vec right to left). */
static inline float rng_phi(orc_rng* r) { return (2.0f * PI_F) * orc_rng_unit(r); }
static inline orc_vec3 rng_uniform_vector(orc_rng* r)
{
	const float phi = rng_phi(r);
	const float ct = 2.0f * orc_rng_unit(r) - 1.0f;
	return make_unit_vec(ct, phi);
}
static inline float rng_exponential(orc_rng* r, float tau) { return -tau * m_log(orc_rng_unit(r)); }

/* --------------------------------------------------------------- tables -- */
/* common/util/table_1D.inl:66-81 */
float orc_t1_get(const orc_table1d* t, float x)
{
	const float x_index = (x - t->x_min) * t->x_step;
	const int low = clampi(floorr(x_index), 0, t->n - 2);
	const float lv = t->data[low], hv = t->data[low + 1];
	const float frac = x_index - (float)low;
	return (1.0f - frac) * lv + frac * hv;
}

#define T2(t, i, j) ((t)->data[(size_t)(i) * (size_t)(t)->ny + (size_t)(j)])
#define T3(t, i, j, k) ((t)->data[((size_t)(j) + (size_t)(i) * (size_t)(t)->ny) * (size_t)(t)->nz + (size_t)(k)])

/* common/util/table_2D.inl:73-94 */
float orc_t2_get(const orc_table2d* t, float x, float y)
{
	const float xi = (x - t->x_min) * t->x_step;
	const float yi = (y - t->y_min) * t->y_step;
	const int lx = clampi(floorr(xi), 0, t->nx - 2);
	const int ly = clampi(floorr(yi), 0, t->ny - 2);
	const float v00 = T2(t, lx, ly), v10 = T2(t, lx + 1, ly);
	const float v01 = T2(t, lx, ly + 1), v11 = T2(t, lx + 1, ly + 1);
	const float fx = xi - (float)lx, fy = yi - (float)ly;
	return (1.0f - fx) * (1.0f - fy) * v00
		+ fx * (1.0f - fy) * v10
		+ (1.0f - fx) * fy * v01
		+ fx * fy * v11;
}

/* common/util/table_2D.inl:96-103 */
float orc_t2_get_rounddown(const orc_table2d* t, float x, float y)
{
	const float xi = (x - t->x_min) * t->x_step;
	const float yi = (y - t->y_min) * t->y_step;
	const int lx = clampi(floorr(xi), 0, t->nx - 1);
	const int ly = clampi(floorr(yi), 0, t->ny - 1);
	return T2(t, lx, ly);
}

/* common/util/table_2D.inl:105-114 */
float orc_t2_get_nearest(const orc_table2d* t, float x, float y)
{
	const float xi = (x - t->x_min) * t->x_step;
	const float yi = (y - t->y_min) * t->y_step;
	const int nx = clampi(rintr(xi), 0, t->nx - 1);
	const int ny = clampi(rintr(yi), 0, t->ny - 1);
	return T2(t, nx, ny);
}

/* common/util/table_3D.inl:78-109 */
float orc_t3_get(const orc_table3d* t, float x, float y, float z)
{
	const float xi = (x - t->x_min) * t->x_step;
	const float yi = (y - t->y_min) * t->y_step;
	const float zi = (z - t->z_min) * t->z_step;
	const int lx = clampi(floorr(xi), 0, t->nx - 2);
	const int ly = clampi(floorr(yi), 0, t->ny - 2);
	const int lz = clampi(floorr(zi), 0, t->nz - 2);
	const float v000 = T3(t, lx, ly, lz), v100 = T3(t, lx + 1, ly, lz);
	const float v010 = T3(t, lx, ly + 1, lz), v110 = T3(t, lx + 1, ly + 1, lz);
	const float v001 = T3(t, lx, ly, lz + 1), v101 = T3(t, lx + 1, ly, lz + 1);
	const float v011 = T3(t, lx, ly + 1, lz + 1), v111 = T3(t, lx + 1, ly + 1, lz + 1);
	const float fx = xi - (float)lx, fy = yi - (float)ly, fz = zi - (float)lz;
	return (1.0f - fx) * (1.0f - fy) * (1.0f - fz) * v000
		+ fx * (1.0f - fy) * (1.0f - fz) * v100
		+ (1.0f - fx) * fy * (1.0f - fz) * v010
		+ fx * fy * (1.0f - fz) * v110
		+ (1.0f - fx) * (1.0f - fy) * fz * v001
		+ fx * (1.0f - fy) * fz * v101
		+ (1.0f - fx) * fy * fz * v011
		+ fx * fy * fz * v111;
}

/* common/util/table_3D.inl:111-121 */
float orc_t3_get_rounddown(const orc_table3d* t, float x, float y, float z)
{
	const float xi = (x - t->x_min) * t->x_step;
	const float yi = (y - t->y_min) * t->y_step;
	const float zi = (z - t->z_min) * t->z_step;
	const int lx = clampi(floorr(xi), 0, t->nx - 1);
	const int ly = clampi(floorr(yi), 0, t->ny - 1);
	const int lz = clampi(floorr(zi), 0, t->nz - 1);
	return T3(t, lx, ly, lz);
}

void orc_t1_probe(const orc_table1d* t, const float* x, float* out, size_t n)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = orc_t1_get(t, x[i]);
}

void orc_t2_probe(const orc_table2d* t, int mode, const float* x, const float* y, float* out, size_t n)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = mode == 0 ? orc_t2_get(t, x[i], y[i])
			: mode == 1 ? orc_t2_get_rounddown(t, x[i], y[i])
			: orc_t2_get_nearest(t, x[i], y[i]);
}

void orc_t3_probe(const orc_table3d* t, int mode, const float* x, const float* y, const float* z, float* out, size_t n)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = mode == 0 ? orc_t3_get(t, x[i], y[i], z[i]) : orc_t3_get_rounddown(t, x[i], y[i], z[i]);
}

/* ------------------------------------------------------------- triangle -- */
void orc_make_triangle(const float r0[3], const float r1[3], const float r2[3],
	int32_t mat_in, int32_t mat_out, orc_triangle* out)
{
	/* core/triangle.inl:1-4: e1 = r1 - r0, e2 = r2 - r0 in float. */
	out->mat_in = mat_in;
	out->mat_out = mat_out;
	out->r0.x = r0[0]; out->r0.y = r0[1]; out->r0.z = r0[2];
	out->e1.x = r1[0] - r0[0]; out->e1.y = r1[1] - r0[1]; out->e1.z = r1[2] - r0[2];
	out->e2.x = r2[0] - r0[0]; out->e2.y = r2[1] - r0[1]; out->e2.z = r2[2] - r0[2];
}

/* core/triangle.inl:7-28 (Moller-Trumbore with absolute tolerance). */
float orc_intersect_ray(const orc_triangle* t, orc_vec3 start, orc_vec3 dir)
{
	const orc_vec3 pvec = cross(dir, t->e2);
	const float det = dot(t->e1, pvec);
	if (fabsf(det) < EPS)
		return -1.0f;
	const orc_vec3 s = v_sub(start, t->r0);
	const float u = dot(s, pvec) / det;
	if ((u < -EPS) || (u > 1.0f + EPS))
		return -1.0f;
	const orc_vec3 qvec = cross(s, t->e1);
	const float v = dot(dir, qvec) / det;
	if ((v < -EPS) || (u + v > 1.0f + EPS))
		return -1.0f;
	const float tt = dot(t->e2, qvec) / det;
	if (tt < 0.0f)
		return -1.0f;
	return tt;
}

void orc_tri_intersect_batch(const orc_triangle* tris, const int32_t* tri_idx,
	const float* pos, const float* dir, float* out_t, size_t n)
{
	for (size_t i = 0; i < n; ++i) {
		orc_vec3 p = { pos[3 * i], pos[3 * i + 1], pos[3 * i + 2] };
		orc_vec3 d = { dir[3 * i], dir[3 * i + 1], dir[3 * i + 2] };
		out_t[i] = orc_intersect_ray(&tris[tri_idx[i]], p, d);
	}
}

/* ------------------------------------------------------------ geometry -- */
/* geometry/octree.inl:24-30 */
int orc_in_domain(const orc_geometry* g, orc_vec3 pos)
{
	const orc_vec3 c = g->center, h = g->halfsize;
	return (pos.x > c.x - h.x) && (pos.x < c.x + h.x)
		&& (pos.y > c.y - h.y) && (pos.y < c.y + h.y)
		&& (pos.z > c.z - h.z) && (pos.z < c.z + h.z);
}

static inline int clz64(uint64_t x) { return __builtin_clzll(x); }

/* geometry/octree.inl:32-192 */
int32_t orc_propagate(const orc_geometry* g, orc_vec3 start, orc_vec3 dir,
	float distance, int32_t ignore_tri, int32_t ignore_mat, float* out_dist,
	uint32_t* out_cells)
{
	const int32_t* od = g->octree_data;
	const orc_vec3 hs = g->halfsize;
	orc_vec3 cur = start;
	float to_go = distance;
	uint64_t location = 1;
	uint32_t cells = 0;
	do {
		float ax = g->center.x, ay = g->center.y, az = g->center.z, aw = 1.0f;
		int index = 0;
		/* re-descend from the root along the location code (:62-69) */
		for (int i = 60 - clz64(location); i >= 0; i -= 3) {
			const unsigned oct = (unsigned)(location >> i) & 7u;
			index = od[index + (int)oct];
			ax += aw * hs.x * ((float)(oct & 1u) - 0.5f);
			ay += aw * hs.y * ((float)((oct & 2u) >> 1) - 0.5f);
			az += aw * hs.z * ((float)((oct & 4u) >> 2) - 0.5f);
			aw *= 0.5f;
		}
		/* descend to the leaf containing the position (:72-83) */
		while (index >= 0) {
			unsigned oct = 0;
			oct ^= (cur.x > ax) ? 1u : 0u;
			oct ^= (cur.y > ay) ? 2u : 0u;
			oct ^= (cur.z > az) ? 4u : 0u;
			location = (location << 3) | oct;
			index = od[index + (int)oct];
			ax += aw * hs.x * ((float)(oct & 1u) - 0.5f);
			ay += aw * hs.y * ((float)((oct & 2u) >> 1) - 0.5f);
			az += aw * hs.z * ((float)((oct & 4u) >> 2) - 0.5f);
			aw *= 0.5f;
		}
		index = -index;
		++cells;

		/* exit plane of the leaf (:98-112 with AABB_intersect :211-220) */
		float isect;
		int32_t target;
		{
			const float hx = aw * hs.x, hy = aw * hs.y, hz = aw * hs.z;
			const float tx = (ax + copysignf(hx + EPS, dir.x) - cur.x) / dir.x;
			const float ty = (ay + copysignf(hy + EPS, dir.y) - cur.y) / dir.y;
			const float tz = (az + copysignf(hz + EPS, dir.z) - cur.z) / dir.z;
			if ((tx < ty) && (tx < tz)) { isect = tx; target = -1; }
			else if ((ty < tx) && (ty < tz)) { isect = ty; target = -2; }
			else { isect = tz; target = -4; }
		}

		/* leaf triangles (:116-147) */
		for (;;) {
			const int32_t ti = od[index++];
			if (ti < 0)
				break;
			if (ti == ignore_tri)
				continue;
			const orc_triangle* tri = &g->triangles[ti];
			const int32_t mat_out = (dot(cross(tri->e1, tri->e2), dir) < 0.0f) ? tri->mat_in : tri->mat_out;
			if (mat_out == ignore_mat)
				continue;
			const float t = orc_intersect_ray(tri, cur, dir);
			if ((t > 0.0f) && (t <= isect + EPS)) {
				isect = t;
				target = ti;
			}
		}

		if (isect >= to_go) {
			*out_dist = distance;
			if (out_cells) *out_cells = cells;
			return -1;
		} else if (target >= 0) {
			*out_dist = distance - to_go + isect;
			if (out_cells) *out_cells = cells;
			return target;
		}

		to_go -= isect;
		cur = v_add(cur, v_mul(dir, isect));

		/* neighbour by location-code bit tricks (:160-186) */
		const unsigned mask = (unsigned)(-target);
		unsigned value;
		if (mask == 1u) value = (dir.x >= 0.0f) ? 0u : 1u;
		else if (mask == 2u) value = (dir.y >= 0.0f) ? 0u : 2u;
		else value = (dir.z >= 0.0f) ? 0u : 4u;
		while (location > 1) {
			if ((location & (uint64_t)mask) == (uint64_t)value) {
				location ^= (uint64_t)mask;
				break;
			}
			location >>= 3;
		}
	} while (location > 1 && cells < ORC_MAX_CELLS);
	*out_dist = distance;
	if (out_cells) *out_cells = cells;
	return -1;
}

/* geometry/trilist.inl:52-85: brute force over all triangles (strict <, no
 * tolerance on the running best).  An independent check of propagate. */
int32_t orc_trilist_propagate(const orc_triangle* tris, int32_t n, orc_vec3 start,
	orc_vec3 dir, float distance, int32_t ignore_tri, int32_t ignore_mat, float* out_dist)
{
	float best = distance;
	int32_t hit = -1;
	for (int32_t i = 0; i < n; ++i) {
		if (i == ignore_tri)
			continue;
		const orc_triangle* tri = &tris[i];
		const int32_t mat_out = (dot(cross(tri->e1, tri->e2), dir) < 0.0f) ? tri->mat_in : tri->mat_out;
		if (mat_out == ignore_mat)
			continue;
		const float t = orc_intersect_ray(tri, start, dir);
		if (t > 0.0f && t < best) {
			best = t;
			hit = i;
		}
	}
	*out_dist = best;
	return hit;
}

void orc_geometry_queries(const orc_geometry* g, const float* pos, const float* dir,
	const float* dist, const int32_t* ignore_tri, const int32_t* ignore_mat,
	int32_t* out_tri, float* out_dist, size_t n)
{
	for (size_t i = 0; i < n; ++i) {
		orc_vec3 p = { pos[3 * i], pos[3 * i + 1], pos[3 * i + 2] };
		orc_vec3 d = { dir[3 * i], dir[3 * i + 1], dir[3 * i + 2] };
		out_tri[i] = orc_propagate(g, p, d, dist[i], ignore_tri[i], ignore_mat[i], &out_dist[i], NULL);
	}
}

/* ------------------------------------------------------ octree builder -- */
/* Triangle-box overlap by the separating-axis test of T. Akenine-Moller,
 * "Fast 3D triangle-box overlap testing" (JGT 2001), in double precision, as
 * used by geometry/octree/tribox.cc:109-184.  Each axis test uses the same
 * vertex pair and product order as there, so borderline cases agree. */
static int axis_test(double a, double b, double fa, double fb,
	const double* va, const double* vb, int i0, int i1, int sign_first,
	double h0, double h1)
{
	double pa, pb;
	if (sign_first) { /* -a*v[i0] + b*v[i1] */
		pa = -a * va[i0] + b * va[i1];
		pb = -a * vb[i0] + b * vb[i1];
	} else {          /* a*v[i0] - b*v[i1] */
		pa = a * va[i0] - b * va[i1];
		pb = a * vb[i0] - b * vb[i1];
	}
	const double mn = pa < pb ? pa : pb;
	const double mx = pa < pb ? pb : pa;
	const double rad = fa * h0 + fb * h1;
	return !(mn > rad || mx < -rad);
}

int orc_tribox_overlap(const double c[3], const double h[3], const double verts[9])
{
	double v0[3], v1[3], v2[3], e0[3], e1[3], e2[3], n[3];
	for (int q = 0; q < 3; ++q) {
		v0[q] = verts[q] - c[q];
		v1[q] = verts[3 + q] - c[q];
		v2[q] = verts[6 + q] - c[q];
	}
	for (int q = 0; q < 3; ++q) {
		e0[q] = v1[q] - v0[q];
		e1[q] = v2[q] - v1[q];
		e2[q] = v0[q] - v2[q];
	}
	/* bullet 1: axis-aligned extents */
	for (int q = 0; q < 3; ++q) {
		double mn = v0[q], mx = v0[q];
		if (v1[q] < mn) mn = v1[q];
		if (v1[q] > mx) mx = v1[q];
		if (v2[q] < mn) mn = v2[q];
		if (v2[q] > mx) mx = v2[q];
		if (mn > h[q] || mx < -h[q])
			return 0;
	}
	/* bullet 2: plane of the triangle against the box */
	n[0] = e0[1] * e1[2] - e0[2] * e1[1];
	n[1] = e0[2] * e1[0] - e0[0] * e1[2];
	n[2] = e0[0] * e1[1] - e0[1] * e1[0];
	{
		double vmin[3], vmax[3];
		for (int q = 0; q < 3; ++q) {
			if (n[q] > 0) { vmin[q] = -h[q] - v0[q]; vmax[q] = h[q] - v0[q]; }
			else { vmin[q] = h[q] - v0[q]; vmax[q] = -h[q] - v0[q]; }
		}
		if (n[0] * vmin[0] + n[1] * vmin[1] + n[2] * vmin[2] > 0)
			return 0;
		if (!(n[0] * vmax[0] + n[1] * vmax[1] + n[2] * vmax[2] >= 0))
			return 0;
	}
	/* bullet 3: the nine edge-cross-axis tests */
	const double* E[3] = { e0, e1, e2 };
	/* vertex pairs per (edge, axis): see the comment above */
	const double* px[3][2] = { { v0, v2 }, { v0, v2 }, { v0, v1 } };
	const double* py[3][2] = { { v0, v2 }, { v0, v2 }, { v0, v1 } };
	const double* pz[3][2] = { { v1, v2 }, { v0, v1 }, { v1, v2 } };
	for (int k = 0; k < 3; ++k) {
		const double* e = E[k];
		const double fex = fabs(e[0]), fey = fabs(e[1]), fez = fabs(e[2]);
		if (!axis_test(e[2], e[1], fez, fey, px[k][0], px[k][1], 1, 2, 0, h[1], h[2])) return 0;
		if (!axis_test(e[2], e[0], fez, fex, py[k][0], py[k][1], 0, 2, 1, h[0], h[2])) return 0;
		if (!axis_test(e[1], e[0], fey, fex, pz[k][0], pz[k][1], 0, 1, 0, h[0], h[1])) return 0;
	}
	return 1;
}

static orc_vec3 tri_r1(const orc_triangle* t) { return v_add(t->r0, t->e1); }
static orc_vec3 tri_r2(const orc_triangle* t) { return v_add(t->r0, t->e2); }

/* octree_builder.cpp:30-39: box inflated by 1.01 in double. */
static int node_overlap(const orc_triangle* t, orc_vec3 c, orc_vec3 h)
{
	const double center[3] = { c.x, c.y, c.z };
	const double half[3] = { 1.01 * (double)h.x, 1.01 * (double)h.y, 1.01 * (double)h.z };
	const orc_vec3 a = t->r0, b = tri_r1(t), d = tri_r2(t);
	const double v[9] = { a.x, a.y, a.z, b.x, b.y, b.z, d.x, d.y, d.z };
	return orc_tribox_overlap(center, half, v) > 0;
}

static int veq(orc_vec3 a, orc_vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

/* core/triangle.inl:86-103 */
static int perfect_overlap(const orc_triangle* a, const orc_triangle* b)
{
	const orc_vec3 a0 = a->r0, a1 = tri_r1(a), a2 = tri_r2(a);
	const orc_vec3 b0 = b->r0, b1 = tri_r1(b), b2 = tri_r2(b);
	return (veq(a0, b0) && veq(a1, b1) && veq(a2, b2))
		|| (veq(a0, b1) && veq(a1, b2) && veq(a2, b0))
		|| (veq(a0, b2) && veq(a1, b0) && veq(a2, b1))
		|| (veq(a0, b0) && veq(a1, b2) && veq(a2, b1))
		|| (veq(a0, b1) && veq(a1, b0) && veq(a2, b2))
		|| (veq(a0, b2) && veq(a1, b1) && veq(a2, b0));
}

typedef struct {
	int leaf;
	int child[8];
	int32_t* tris;
	int ntris;
} bnode;

typedef struct {
	bnode* nodes;
	int n, cap;
	const orc_triangle* tris;
	int failed;
} bctx;

static int new_node(bctx* b)
{
	if (b->n == b->cap) {
		int nc = b->cap ? 2 * b->cap : 64;
		bnode* nn = (bnode*)realloc(b->nodes, (size_t)nc * sizeof(bnode));
		if (!nn) { b->failed = 1; return -1; }
		b->nodes = nn;
		b->cap = nc;
	}
	memset(&b->nodes[b->n], 0, sizeof(bnode));
	return b->n++;
}

/* octree_builder.cpp:48-65 */
static void octant_box(int oct, orc_vec3 c, orc_vec3 h, orc_vec3* oc, orc_vec3* oh)
{
	const orc_vec3 hh = s_mul(0.5f, h);
	*oh = hh;
	oc->x = c.x + (float)((((oct & 1) << 1) - 1)) * hh.x;
	oc->y = c.y + (float)((((oct & 2)) - 1)) * hh.y;
	oc->z = c.z + (float)((((oct & 4) >> 1) - 1)) * hh.z;
}

/* Top-down form of the insertion-driven construction of
 * octree_builder.cpp:93-155: a node at depth d holding the (duplicate-free,
 * index-ordered) triangles that overlap it and all its ancestors splits iff it
 * holds more than 16 and d < 21.  This gives the same tree as inserting the
 * triangles one by one, because a node's set only grows during insertion and
 * leaf lists stay in insertion (= index) order. */
static void build_node(bctx* b, int id, orc_vec3 c, orc_vec3 h, int depth,
	const int32_t* cand, int ncand)
{
	/* overlap with this node and leaf-level duplicate rejection
	 * (octree_builder.cpp:96-115) */
	int32_t* mine = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncand > 0 ? ncand : 1));
	int nm = 0;
	if (!mine) { b->failed = 1; return; }
	for (int i = 0; i < ncand; ++i) {
		const orc_triangle* t = &b->tris[cand[i]];
		if (!node_overlap(t, c, h))
			continue;
		int dup = 0;
		for (int j = 0; j < nm && !dup; ++j)
			dup = perfect_overlap(t, &b->tris[mine[j]]);
		if (dup) {
			fprintf(stderr, "WARNING: ignoring duplicate triangle!\n");
			continue;
		}
		mine[nm++] = cand[i];
	}
	if (nm > 16 && depth < 21) {
		b->nodes[id].leaf = 0;
		for (int oct = 0; oct < 8; ++oct) {
			const int cid = new_node(b);
			if (cid < 0) { free(mine); return; }
			b->nodes[id].child[oct] = cid;
		}
		for (int oct = 0; oct < 8; ++oct) {
			orc_vec3 oc, oh;
			octant_box(oct, c, h, &oc, &oh);
			build_node(b, b->nodes[id].child[oct], oc, oh, depth + 1, mine, nm);
		}
		free(mine);
	} else {
		b->nodes[id].leaf = 1;
		b->nodes[id].tris = mine;
		b->nodes[id].ntris = nm;
	}
}

int orc_octree_build(const orc_triangle* tris, int32_t ntri,
	int32_t** octree_data, int32_t* n_octree,
	orc_triangle** sorted_tris, int32_t* n_sorted,
	float center[3], float halfsize[3])
{
	if (ntri <= 0)
		return -1;
	/* root box (octree_builder.cpp:171-209) */
	orc_vec3 mn, mx;
	for (int32_t i = 0; i < ntri; ++i) {
		const orc_triangle* t = &tris[i];
		const orc_vec3 A = t->r0, B = tri_r1(t), C = tri_r2(t);
		const orc_vec3 tmin = { fminf(fminf(A.x, B.x), C.x), fminf(fminf(A.y, B.y), C.y), fminf(fminf(A.z, B.z), C.z) };
		const orc_vec3 tmax = { fmaxf(fmaxf(A.x, B.x), C.x), fmaxf(fmaxf(A.y, B.y), C.y), fmaxf(fmaxf(A.z, B.z), C.z) };
		if (i == 0) { mn = tmin; mx = tmax; }
		else {
			mn.x = fmin_std(mn.x, tmin.x); mn.y = fmin_std(mn.y, tmin.y); mn.z = fmin_std(mn.z, tmin.z);
			mx.x = fmax_std(mx.x, tmax.x); mx.y = fmax_std(mx.y, tmax.y); mx.z = fmax_std(mx.z, tmax.z);
		}
	}
	const orc_vec3 c = s_mul(0.5f, v_add(mn, mx));
	orc_vec3 h = s_mul(0.5f, v_sub(mx, mn));
	h.x += 1.0f; h.y += 1.0f; h.z += 1.0f;

	bctx b;
	memset(&b, 0, sizeof(b));
	b.tris = tris;
	/* the root keeps every triangle that overlaps it (:211-225) and always
	 * has eight children */
	int32_t* all = (int32_t*)malloc(sizeof(int32_t) * (size_t)ntri);
	if (!all) return -2;
	int nall = 0;
	for (int32_t i = 0; i < ntri; ++i)
		if (node_overlap(&tris[i], c, h))
			all[nall++] = i;
	const int root = new_node(&b);
	b.nodes[root].leaf = 0;
	for (int oct = 0; oct < 8; ++oct)
		b.nodes[root].child[oct] = new_node(&b);
	for (int oct = 0; oct < 8; ++oct) {
		orc_vec3 oc, oh;
		octant_box(oct, c, h, &oc, &oh);
		build_node(&b, b.nodes[root].child[oct], oc, oh, 1, all, nall);
	}
	free(all);
	if (b.failed) return -2;

	/* breadth-first order (= Morton order) (:252-276) */
	int* order = (int*)malloc(sizeof(int) * (size_t)b.n);
	int head = 0, tail = 0;
	order[tail++] = root;
	while (head < tail) {
		const int id = order[head++];
		if (!b.nodes[id].leaf)
			for (int oct = 0; oct < 8; ++oct)
				order[tail++] = b.nodes[id].child[oct];
	}

	/* triangles by first appearance in a leaf (:284-308) */
	int32_t* map = (int32_t*)malloc(sizeof(int32_t) * (size_t)ntri);
	orc_triangle* sorted = (orc_triangle*)malloc(sizeof(orc_triangle) * (size_t)ntri);
	int32_t ns = 0;
	for (int32_t i = 0; i < ntri; ++i) map[i] = -1;
	size_t total = 0;
	for (int k = 0; k < tail; ++k) {
		const bnode* nd = &b.nodes[order[k]];
		if (!nd->leaf) { total += 8; continue; }
		total += (size_t)nd->ntris + 1;
		for (int j = 0; j < nd->ntris; ++j) {
			const int32_t ti = nd->tris[j];
			if (map[ti] >= 0) continue;
			map[ti] = ns;
			sorted[ns++] = tris[ti];
		}
	}

	/* index table (:310-347) */
	int32_t* od = (int32_t*)malloc(sizeof(int32_t) * total);
	int* start = (int*)malloc(sizeof(int) * (size_t)b.n);
	size_t pos = 0;
	for (int k = 0; k < tail; ++k) {
		const int id = order[k];
		const bnode* nd = &b.nodes[id];
		start[id] = (int)pos;
		if (nd->leaf) {
			for (int j = 0; j < nd->ntris; ++j)
				od[pos++] = map[nd->tris[j]];
			od[pos++] = -1;
		} else {
			for (int oct = 0; oct < 8; ++oct)
				od[pos++] = 0;
		}
	}
	for (int k = 0; k < tail; ++k) {
		const int id = order[k];
		const bnode* nd = &b.nodes[id];
		if (nd->leaf) continue;
		for (int oct = 0; oct < 8; ++oct) {
			const int cid = nd->child[oct];
			od[start[id] + oct] = b.nodes[cid].leaf ? -start[cid] : start[cid];
		}
	}

	for (int i = 0; i < b.n; ++i)
		free(b.nodes[i].tris);
	free(b.nodes);
	free(order);
	free(map);
	free(start);

	*octree_data = od;
	*n_octree = (int32_t)total;
	*sorted_tris = sorted;
	*n_sorted = ns;
	center[0] = c.x; center[1] = c.y; center[2] = c.z;
	halfsize[0] = h.x; halfsize[1] = h.y; halfsize[2] = h.z;
	return 0;
}

void orc_free(void* p) { free(p); }

/* ------------------------------------------------- material preprocessing -- */
/* common/util/range.h:264-323: geomspace value i = first * exp(i * step),
 * step = log(stop/start)/(n-1). */
static double geom_at(double front, double back, int n, int i)
{
	const double step = log(back / front) / (double)(size_t)(n - 1);
	return front * exp((double)(size_t)i * step);
}
/* range.h:210-262: linspace value i = first + i * step. */
static double lin_at(double front, double back, int n, int i)
{
	const double step = (back - front) / (double)(size_t)(n - 1);
	return front + (double)(size_t)i * step;
}

static float* copy_data(const orc_raw_dataset* d, size_t count)
{
	float* p = (float*)malloc(sizeof(float) * (count ? count : 1));
	if (p && count) memcpy(p, d->data, sizeof(float) * count);
	return p;
}

/* table set_scale (table_1D.inl:104-109 etc.): step = (n-1)/(max-min) in
 * float. */
static void scale_log_axis(const orc_raw_dataset* d, int ax, int n, float* mn, float* step)
{
	const float lo = (float)log(d->ax_front[ax]);
	const float hi = (float)log(geom_at(d->ax_front[ax], d->ax_back[ax], n, n - 1));
	*mn = lo;
	*step = (float)(n - 1) / (hi - lo);
}
static void lin_axis(const orc_raw_dataset* d, int ax, int n, double* front, double* back)
{
	if (d->ax_present[ax]) { *front = d->ax_front[ax]; *back = lin_at(d->ax_front[ax], d->ax_back[ax], n, n - 1); }
	else { *front = 0.0; *back = lin_at(0.0, 1.0, n, n - 1); }
}
static void scale_lin_axis(const orc_raw_dataset* d, int ax, int n, float* mn, float* step)
{
	double f, b;
	lin_axis(d, ax, n, &f, &b);
	const float lo = (float)f, hi = (float)b;
	*mn = lo;
	*step = (float)(n - 1) / (hi - lo);
}

/* log(imfp) tables (elastic.h:128-141, full_penn.h:305-318,
 * inelastic.h:222-235). */
static int make_log_imfp(const orc_raw_dataset* d, orc_table1d* t)
{
	if (!d->data || d->ndim != 1) return -1;
	const int n = d->dims[0];
	float* p = copy_data(d, (size_t)n);
	if (!p) return -2;
	const float unit = (float)(d->unit * 1.0);
	for (int x = 0; x < n; ++x)
		p[x] = logf(p[x] * unit);
	t->data = p; t->n = n;
	scale_log_axis(d, 0, n, &t->x_min, &t->x_step);
	return 0;
}

/* log(omega) / log(omega') tables: clamp to K - fermi, then log in double
 * (full_penn.h:320-345, inelastic.h:237-264). */
static int make_log_omega(const orc_raw_dataset* d, double fermi, orc_table2d* t)
{
	if (!d->data || d->ndim != 2) return -1;
	const int nx = d->dims[0], ny = d->dims[1];
	float* p = copy_data(d, (size_t)nx * (size_t)ny);
	if (!p) return -2;
	const float unit = (float)(d->unit / 1.0);
	for (int x = 0; x < nx; ++x) {
		const double K = geom_at(d->ax_front[0], d->ax_back[0], nx, x);
		for (int y = 0; y < ny; ++y) {
			const double v = (double)(p[(size_t)x * ny + y] * unit);
			const double a = (K - fermi) / 1.0;
			const double m = (v < a) ? v : a;           /* std::min<double> */
			const double c = (0.0 < m) ? m : 0.0;       /* std::max(0.0, .) */
			p[(size_t)x * ny + y] = (float)log(c);
		}
	}
	t->data = p; t->nx = nx; t->ny = ny;
	scale_log_axis(d, 0, nx, &t->x_min, &t->x_step);
	scale_lin_axis(d, 1, ny, &t->y_min, &t->y_step);
	return 0;
}

int orc_material_create(const orc_raw_material* r, orc_material* m)
{
	memset(m, 0, sizeof(*m));
	int rc;
	/* core/material.inl:13-18 */
	m->barrier = (float)(r->barrier / 1.0);
	m->el_scale = 1.0f;
	m->in_scale = 1.0f;
	m->inel_model = r->inel_model;

	/* kieft_elastic::create (elastic.h:120-166) */
	if ((rc = make_log_imfp(&r->el_imfp, &m->el_imfp)) != 0) return rc;
	{
		const orc_raw_dataset* d = &r->el_icdf;
		if (!d->data || d->ndim != 2) return -1;
		const int nx = d->dims[0], ny = d->dims[1];
		float* p = copy_data(d, (size_t)nx * (size_t)ny);
		if (!p) return -2;
		for (size_t i = 0; i < (size_t)nx * ny; ++i) {
			const float v = p[i];
			const float a = (v < 1.0f) ? v : 1.0f;      /* std::min(1, v) */
			p[i] = (-1.0f < a) ? a : -1.0f;             /* std::max(-1, .) */
		}
		m->el_icdf.data = p; m->el_icdf.nx = nx; m->el_icdf.ny = ny;
		scale_log_axis(d, 0, nx, &m->el_icdf.x_min, &m->el_icdf.x_step);
		scale_lin_axis(d, 1, ny, &m->el_icdf.y_min, &m->el_icdf.y_step);
	}
	m->phonon_loss = (float)(r->phonon_loss / 1.0);
	m->recoil_const = 2.0f * (float)((2.0 * (9.109383e-28 * 6.2415e15)) / r->effective_A);

	m->fermi = (float)(r->fermi / 1.0);
	m->band_gap = (float)((r->has_band_gap ? r->band_gap : -1.0) / 1.0);

	if (r->inel_model == ORC_INEL_PENN) {
		/* full_penn::create (full_penn.h:294-406) */
		if ((rc = make_log_imfp(&r->in_imfp, &m->in_imfp)) != 0) return rc;
		if ((rc = make_log_omega(&r->in_icdf, r->fermi, &m->in_icdf)) != 0) return rc;
		{
			const orc_raw_dataset* d = &r->pn_q;
			if (!d->data || d->ndim != 3) return -1;
			const int nx = d->dims[0], ny = d->dims[1], nz = d->dims[2];
			const size_t cnt = (size_t)nx * ny * nz;
			float* p = copy_data(d, cnt);
			if (!p) return -2;
			const float unit = (float)(d->unit * 0.19519 * 1.0);
			for (size_t i = 0; i < cnt; ++i)
				p[i] = p[i] * unit;
			orc_table3d* t = &m->pn_q;
			t->data = p; t->nx = nx; t->ny = ny; t->nz = nz;
			scale_log_axis(d, 0, nx, &t->x_min, &t->x_step);
			scale_lin_axis(d, 1, ny, &t->y_min, &t->y_step);
			scale_lin_axis(d, 2, nz, &t->z_min, &t->z_step);
		}
		{
			const orc_raw_dataset* d = &r->pn_ion;
			if (!d->data || d->ndim != 3) return -1;
			const int nx = d->dims[0], ny = d->dims[1], nz = d->dims[2];
			const size_t cnt = (size_t)nx * ny * nz;
			float* p = copy_data(d, cnt);
			if (!p) return -2;
			const float unit = (float)(d->unit / 1.0);
			for (size_t i = 0; i < cnt; ++i) {
				float b = p[i] * unit;
				if (b < 50.0f || !isfinite(b))
					b = -1.0f;
				p[i] = b;
			}
			orc_table3d* t = &m->pn_ion;
			t->data = p; t->nx = nx; t->ny = ny; t->nz = nz;
			scale_log_axis(d, 0, nx, &t->x_min, &t->x_step);
			scale_log_axis(d, 1, ny, &t->y_min, &t->y_step);
			scale_lin_axis(d, 2, nz, &t->z_min, &t->z_step);
		}
	} else {
		/* kieft_inelastic::create (inelastic.h:212-327) */
		if ((rc = make_log_imfp(&r->in_imfp, &m->in_imfp)) != 0) return rc;
		if ((rc = make_log_omega(&r->in_icdf, r->fermi, &m->in_icdf)) != 0) return rc;
		const orc_raw_dataset* d = &r->ki_ion;
		if (!d->data || d->ndim != 2) return -1;
		const int nx = d->dims[0], ny = d->dims[1];
		float* p = copy_data(d, (size_t)nx * (size_t)ny);
		if (!p) return -2;
		orc_table2d* t = &m->ki_ion;
		t->data = p; t->nx = nx; t->ny = ny;
		scale_log_axis(d, 0, nx, &t->x_min, &t->x_step);
		scale_lin_axis(d, 1, ny, &t->y_min, &t->y_step);
		/* outer shells below 100 eV, sorted ascending (:271-285) */
		double* shells = (double*)malloc(sizeof(double) * (size_t)(r->n_outer_shells + 1));
		int ns = 0;
		for (int i = 0; i < r->n_outer_shells; ++i) {
			const double v = r->outer_shells[i] * r->outer_shells_unit;
			if (v < 100.0)
				shells[ns++] = v;
		}
		for (int i = 1; i < ns; ++i)
			for (int j = i; j > 0 && shells[j] < shells[j - 1]; --j) {
				const double tmp = shells[j]; shells[j] = shells[j - 1]; shells[j - 1] = tmp;
			}
		/* in-place rebuild that reads its own partly-rewritten rows
		 * (:287-321) */
		double Pf, Pb;
		lin_axis(d, 1, ny, &Pf, &Pb);
		(void)Pb;
		const double unit = d->unit;
		for (int x = 0; x < nx; ++x) {
			const double K = geom_at(d->ax_front[0], d->ax_back[0], nx, x);
			for (int y = 0; y < ny; ++y) {
				const float P = (float)lin_at(Pf, d->ax_present[1] ? d->ax_back[1] : 1.0, ny, y);
				double binding = -1.0;
				if (K > 100.0) {
					binding = (double)orc_t2_get_rounddown(t, (float)log((K + 10.0) / 1.0), P) * unit;
					if (binding < 50.0 || !isfinite(binding))
						binding = -1.0;
				}
				if (binding < 0.0) {
					/* largest outer shell <= K */
					for (int s = ns - 1; s >= 0; --s)
						if (!(shells[s] > K)) { binding = shells[s]; break; }
				}
				p[(size_t)x * ny + y] = (float)(binding / 1.0);
			}
		}
		free(shells);
	}
	return 0;
}

void orc_material_free(orc_material* m)
{
	free((void*)m->el_imfp.data);
	free((void*)m->el_icdf.data);
	free((void*)m->in_imfp.data);
	free((void*)m->in_icdf.data);
	free((void*)m->pn_q.data);
	free((void*)m->pn_ion.data);
	free((void*)m->ki_ion.data);
	memset(m, 0, sizeof(*m));
}

/* scatter_list::get_max_energy (core/scatter_list.h:59-64, elastic.h:46-49). */
float orc_material_max_energy(const orc_material* m)
{
	float e = INFINITY;
	const float a = expf(m->in_imfp.x_min + (float)(m->in_imfp.n - 1) / m->in_imfp.x_step);
	const float b = expf(m->el_imfp.x_min + (float)(m->el_imfp.n - 1) / m->el_imfp.x_step);
	if (b < e) e = b;
	if (a < e) e = a;
	return e;
}

/* -------------------------------------------------------------- physics -- */
/* full_penn.h:50-57 / elastic.h:54-62 / inelastic.h:54-61; the mean free
 * path is multiplied by the mechanism's scale (exact when it is 1).  *tau
 * receives the scaled mean free path. */
static float mech_sample_path(const orc_table1d* t, float scale, const orc_particle* p, orc_rng* rng,
	float* tau)
{
	const float imfp = m_exp(orc_t1_get(t, m_log(p->E)));
	*tau = (1.0f / imfp) * scale;
	return rng_exponential(rng, *tau);
}

/* core/scatter_list.h:101-107 + common/tuple.h:114-124: the elastic
 * mechanism (index 1) is sampled before the inelastic one (index 0); the
 * strictly shorter path wins. */
static void sample_path_tau(const orc_material* m, const orc_particle* p, orc_rng* rng,
	uint8_t* type, float* distance, float* tau_el, float* tau_in)
{
	uint8_t ty = 0;
	float dist = INFINITY;
	const float pe = mech_sample_path(&m->el_imfp, m->el_scale, p, rng, tau_el);
	if (pe < dist) { ty = 2; dist = pe; }
	const float pi = mech_sample_path(&m->in_imfp, m->in_scale, p, rng, tau_in);
	if (pi < dist) { ty = 1; dist = pi; }
	*type = ty;
	*distance = dist;
}

void orc_sample_path(const orc_material* m, const orc_particle* p, orc_rng* rng,
	uint8_t* type, float* distance)
{
	float te, ti;
	sample_path_tau(m, p, rng, type, distance, &te, &ti);
}

/* physics/kieft/elastic.h:67-115 */
void orc_elastic_execute(const orc_material* m, const orc_physics_opts* o,
	orc_particle* p, orc_rng* rng)
{
	orc_particle tp = *p;
	float cos_theta, sin_theta;
	{
		const float x = m_log(p->E);
		const float y = orc_rng_unit(rng);
		cos_theta = clampr(orc_t2_get(&m->el_icdf, x, y), -1.0f, 1.0f);
		sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
	}
	tp.dir = normalised(tp.dir);
	const orc_vec3 normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
	tp.dir = v_add(v_mul(tp.dir, cos_theta), v_mul(normal_dir, sin_theta));
	if (tp.E < 200.0f) {
		if (o->el_phonon_loss)
			tp.E -= fmin_std(50e-3f, m->phonon_loss);
	} else {
		if (o->el_atomic_recoil)
			tp.E -= m->recoil_const * (1.0f - cos_theta) * tp.E;
	}
	*p = tp;
}

/* physics/full_penn.h:62-288 */
int orc_penn_execute(const orc_material* m, const orc_physics_opts* o,
	orc_particle* p, orc_rng* rng, orc_particle* se)
{
	orc_particle tp = *p;
	float omega;
	{
		const float x = m_log(tp.E);
		const float y = orc_rng_unit(rng);
		omega = m_exp(orc_t2_get(&m->in_icdf, x, y));
	}
	float binding;
	{
		const float x = m_log(tp.E);
		const float y = m_log(omega / tp.E);
		const float z = orc_rng_unit(rng);
		binding = orc_t3_get_rounddown(&m->pn_ion, x, y, z);
	}
	tp.dir = normalised(tp.dir);

	if (binding <= 0.0f) {
		if (omega > m->band_gap) {
			/* Fermi-sea excitation (:100-231) */
			float q;
			{
				const float x = m_log(tp.E);
				const float y = omega / tp.E;
				const float z = orc_rng_unit(rng);
				q = orc_t3_get(&m->pn_q, x, y, z);
			}
			const orc_vec3 prim_normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
			orc_vec3 q_dir, q_normal_dir;
			{
				const float c = clampr((omega + q * q) / (2.0f * sqrtf(tp.E) * q), -1.0f, 1.0f);
				const float s = sqrtf(1.0f - c * c);
				q_dir = normalised(v_add(v_mul(tp.dir, c), v_mul(prim_normal_dir, s)));
				q_normal_dir = normalised(make_normal_vec(q_dir, rng_phi(rng)));
			}
			const float cpf = clampr((2.0f * tp.E - omega - q * q) / (2.0f * sqrtf(tp.E * (tp.E - omega))), -1.0f, 1.0f);
			const float spf = sqrtf(1.0f - cpf * cpf);
			tp.E -= omega;
			tp.dir = v_add(v_mul(tp.dir, cpf), v_mul(prim_normal_dir, spf));
			*p = tp;

			if (o->pn_generate_secondary) {
				orc_particle sp;
				sp.pos = tp.pos;
				const float k_fermi = sqrtf(m->fermi);
				const float k_fermi_omega = sqrtf(m->fermi + omega);
				if (q < k_fermi_omega - k_fermi) {
					/* plasmon decay (:141-194) */
					const float f = m->fermi / omega;
					float x;
					if (f < 0.1f) {
						const float U = m_cbrt(orc_rng_unit(rng));
						x = f * U * U;
					} else {
						const float lower = fmax_std(0.0f, f - 1.0f);
						const float upper = f;
						const float d1 = upper - lower;
						const float d2 = upper * upper - lower * lower;
						const float alpha = d2 / (d1 + d2);
						int rounds = 0;
						do {
							const float U1 = orc_rng_unit(rng);
							const float U2 = orc_rng_unit(rng);
							x = (U1 < alpha) ? sqrtf(d2 * U2 + lower * lower) : d1 * U2 + lower;
						} while (orc_rng_unit(rng) > sqrtf(x * x + x) / (x + 0.5f) && ++rounds < ORC_MAX_REJECT);
					}
					sp.E = x * omega + omega;
					sp.dir = rng_uniform_vector(rng);
				} else if (q < k_fermi_omega + k_fermi) {
					/* single excitation (:195-222) */
					const float ki_parr = (omega - q * q) / (2.0f * q);
					const float kf_parr = ki_parr + q;
					const float ki_perp_min_sq = fmax_std(0.0f, m->fermi - kf_parr * kf_parr);
					const float ki_perp_max_sq = m->fermi - ki_parr * ki_parr;
					const float ki_perp_sq = fmax_std(0.0f,
						ki_perp_min_sq + (ki_perp_max_sq - ki_perp_min_sq) * orc_rng_unit(rng));
					const orc_vec3 ki = v_add(s_mul(ki_parr, q_dir), s_mul(sqrtf(ki_perp_sq), q_normal_dir));
					sp.E = mag2(ki) + omega;
					sp.dir = v_add(ki, s_mul(q, q_dir));
				} else {
					return 0;
				}
				*se = sp;
				return 1;
			}
			return 0;
		}
		/* sub-band-gap loss (:234-242) */
		tp.E -= omega;
		*p = tp;
		return 0;
	}

	/* inner-shell excitation (:245-287) */
	const orc_vec3 normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
	const float K = tp.E - m->fermi + 2.0f * binding;
	const float dK = binding + omega;
	const float cos_theta = sqrtf(dK / K);
	const float sin_theta = sqrtf(1.0f - saturater(cos_theta * cos_theta));
	orc_vec3 sdir = v_add(v_mul(tp.dir, cos_theta), v_mul(normal_dir, sin_theta));
	sdir = normalised(sdir);
	sdir = v_add(sdir, s_mul(sqrtf(binding / dK), rng_uniform_vector(rng)));
	sdir = normalised(sdir);
	int made = 0;
	if (o->pn_generate_secondary) {
		se->E = m->fermi + omega - binding;
		se->pos = tp.pos;
		se->dir = sdir;
		made = 1;
	}
	tp.E -= omega;
	tp.dir = v_sub(tp.dir, s_mul(cos_theta, sdir));
	*p = tp;
	return made;
}

/* physics/kieft/inelastic.h:68-206 */
int orc_kieft_execute(const orc_material* m, const orc_physics_opts* o,
	orc_particle* p, orc_rng* rng, orc_particle* se)
{
	orc_particle tp = *p;
	float omega0;
	{
		const float x = m_log(tp.E);
		const float y = orc_rng_unit(rng);
		omega0 = m_exp(orc_t2_get(&m->in_icdf, x, y));
	}
	float binding;
	{
		const float x = m_log(omega0);
		const float y = orc_rng_unit(rng);
		binding = orc_t2_get_rounddown(&m->ki_ion, x, y);
	}
	float omega;
	{
		const float omega_max = 0.5f * (tp.E + omega0 - m->fermi);
		float omega_min = omega0;
		float w0 = fmin_std(omega0 - 1.0f, fmax_std(0.0f, binding) - m->fermi);
		if (tp.E > 2.0f * omega0) {
			omega_min = 0.5f * (tp.E + omega0 - sqrtf(tp.E * (tp.E - 2.0f * omega0)));
			w0 = omega0;
		}
		const float U = orc_rng_unit(rng);
		if ((w0 > 0.0f) && (omega_min > w0) && (omega_min < omega_max)) {
			const float f_min = 1.0f / w0 * m_log((omega_min - w0) / omega_min);
			const float f_max = 1.0f / w0 * m_log((omega_max - w0) / omega_max);
			omega = -w0 / m_expm1(w0 * (f_min * (1.0f - U) + f_max * U));
		} else {
			omega = omega0 * omega_max / (omega0 * (1.0f - U) + omega_max * U);
		}
	}
	if (binding < 0.0f) {
		if (m->band_gap < 0.0f) {
			/* metals: no special case in the reference */
		} else if (omega0 > m->band_gap) {
			binding = m->band_gap;
		} else {
			if (o->ki_optical_phonon) {
				tp.E -= omega0;
				*p = tp;
			}
			return 0;
		}
	}
	binding = fmax_std(0.0f, binding);
	tp.dir = normalised(tp.dir);
	const orc_vec3 normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
	const float K = tp.E - m->fermi + 2.0f * binding;
	const float dK = binding + omega;
	const float cos_theta = sqrtf(dK / K);
	const float sin_theta = sqrtf(1.0f - saturater(cos_theta * cos_theta));
	orc_vec3 sdir = v_add(v_mul(tp.dir, cos_theta), v_mul(normal_dir, sin_theta));
	if (o->ki_instantaneous_momentum) {
		sdir = normalised(sdir);
		sdir = v_add(sdir, s_mul(sqrtf(binding / dK), rng_uniform_vector(rng)));
	}
	sdir = normalised(sdir);
	int made = 0;
	if (o->ki_generate_secondary) {
		se->E = m->fermi + omega - binding;
		se->pos = tp.pos;
		se->dir = sdir;
		made = 1;
	}
	tp.E -= omega;
	if (o->ki_momentum_conservation)
		tp.dir = v_sub(tp.dir, s_mul(cos_theta, sdir));
	*p = tp;
	return made;
}

static void add_barrier_score(const orc_score_map* map, double* score, int32_t mi, int32_t mo, float g)
{
	if (!map || g == 0.0f)
		return;
	if (mo >= 0 && mo < ORC_MAX_MATERIALS && map->pidx[mo][ORC_P_BARRIER] >= 0)
		score[map->pidx[mo][ORC_P_BARRIER]] += (double)g;
	if (mi >= 0 && mi < ORC_MAX_MATERIALS && map->pidx[mi][ORC_P_BARRIER] >= 0)
		score[map->pidx[mi][ORC_P_BARRIER]] -= (double)g;
}

/* physics/boundary_intersect.h:38-155 */
static int boundary_execute_scored(const orc_world* w, const orc_triangle* tri,
	orc_particle* p, int32_t* mat, orc_rng* rng, const orc_score_map* map, double* score)
{
	orc_particle tp = *p;
	const orc_vec3 nd = normalised(tp.dir);
	const orc_vec3 n = normalised(cross(tri->e1, tri->e2));
	const float cos_theta = dot(n, nd);
	int32_t mi, mo;
	if (cos_theta > 0.0f) { mi = tri->mat_in; mo = tri->mat_out; }
	else { mi = tri->mat_out; mo = tri->mat_in; }

	switch (mo) {
	case ORC_DETECTOR:
		return ORC_DETECTED;
	case ORC_DETECTOR_LT50:
		return (tp.E < 50.0f) ? ORC_DETECTED : ORC_ALIVE;
	case ORC_DETECTOR_GE50:
		return (tp.E >= 50.0f) ? ORC_DETECTED : ORC_ALIVE;
	case ORC_TERMINATOR:
		return ORC_TERMINATED;
	case ORC_MIRROR:
		tp.dir = v_sub(nd, v_mul(s_mul(2.0f, n), cos_theta));
		*p = tp;
		return ORC_ALIVE;
	case ORC_NOP:
		return ORC_ALIVE;
	default:
		break;
	}

	float dU = 0.0f;
	if (mi >= 0) dU -= w->materials[mi].barrier;
	if (mo >= 0) dU += w->materials[mo].barrier;

	float g = 0.0f;   /* d log p / d dU of the random choices below */
	if (tp.E * cos_theta * cos_theta + dU > 0.0f) {
		const float ec2 = tp.E * cos_theta * cos_theta;
		const float s = sqrtf(1.0f + dU / ec2);
		const float T = w->opts.bi_quantum_transmission ? 4.0f * s / ((1.0f + s) * (1.0f + s)) : 1.0f;
		if (orc_rng_unit(rng) < T) {
			if (w->opts.bi_quantum_transmission && T < 1.0f)
				add_barrier_score(map, score, mi, mo, (1.0f - s) / (2.0f * s * s * (1.0f + s) * ec2));
			if (w->opts.bi_interface_refraction)
				tp.dir = v_add(v_sub(nd, v_mul(n, cos_theta)), v_mul(s_mul(s, n), cos_theta));
			tp.E += dU;
			*mat = mo;
			*p = tp;
			return ORC_ALIVE;
		}
		if (w->opts.bi_quantum_transmission && T < 1.0f)
			g = 2.0f / (s * dU);
	}
	if (w->opts.bi_interface_absorption) {
		if (dU < 0.0f) {
			const float P = m_exp(1.0f + 0.5f * tp.E / dU);
			const int absorbed = orc_rng_unit(rng) < P;
			if (P < 1.0f)
				g = g + (absorbed ? -0.5f * tp.E / (dU * dU) : P * 0.5f * tp.E / (dU * dU * (1.0f - P)));
			if (absorbed) {
				add_barrier_score(map, score, mi, mo, g);
				return ORC_TERMINATED;
			}
		}
	}
	add_barrier_score(map, score, mi, mo, g);
	tp.dir = v_sub(nd, v_mul(s_mul(2.0f, n), cos_theta));
	*p = tp;
	return ORC_ALIVE;
}

int orc_boundary_execute(const orc_world* w, const orc_triangle* tri,
	orc_particle* p, int32_t* mat, orc_rng* rng)
{
	return boundary_execute_scored(w, tri, p, mat, rng, NULL, NULL);
}

/* --------------------------------------------------------------- a step -- */
/* init (drivers/cpu/cpu_driver.inl:74-126 / drivers/gpu/gpu_driver.inl:
 * 403-480) followed by intersect (:128-139 / :482-503) and scatter
 * (:141-160 / :505-566). */
int orc_step_scored(const orc_world* w, int semantics, orc_state* s, orc_rng* rng,
	orc_state* se, int* se_made, int* event, const orc_score_map* map, double* score)
{
	*se_made = 0;
	*event = ORC_EV_NONE;
	if (!orc_in_domain(&w->geom, s->p.pos))
		return ORC_TERMINATED;

	uint8_t type = 0;
	float distance = w->geom.max_extent;
	float tau_el = 0.0f, tau_in = 0.0f;
	if (s->mat >= 0) {
		const orc_material* m = &w->materials[s->mat];
		if (s->p.E < m->barrier + w->energy_threshold)
			return ORC_TERMINATED;
		sample_path_tau(m, &s->p, rng, &type, &distance, &tau_el, &tau_in);
	}

	const orc_vec3 nd = normalised(s->p.dir);
	float hit_dist;
	const int32_t hit = orc_propagate(&w->geom, s->p.pos, nd, distance, s->last_tri, s->mat, &hit_dist, NULL);
	if (semantics == 1)
		s->p.dir = nd;
	if (hit < 0) {
		s->p.pos = v_add(s->p.pos, v_mul(nd, distance));
		*event = type;
	} else {
		s->p.pos = v_add(s->p.pos, v_mul(nd, hit_dist));
		s->last_tri = hit;
		*event = ORC_EV_INTERSECT;
	}
	if (map && s->mat >= 0 && s->mat < ORC_MAX_MATERIALS && type != 0) {
		/* free-path score: distance travelled d, rates 1/tau */
		const orc_material* m = &w->materials[s->mat];
		const float d = (hit < 0) ? distance : hit_dist;
		const int pe = map->pidx[s->mat][ORC_P_EL_SCALE], pi = map->pidx[s->mat][ORC_P_IN_SCALE];
		if (pe >= 0)
			score[pe] += (double)((d / tau_el - ((*event == ORC_EV_ELASTIC) ? 1.0f : 0.0f)) / m->el_scale);
		if (pi >= 0)
			score[pi] += (double)((d / tau_in - ((*event == ORC_EV_INELASTIC) ? 1.0f : 0.0f)) / m->in_scale);
	}

	if (*event == ORC_EV_INTERSECT)
		return boundary_execute_scored(w, &w->geom.triangles[hit], &s->p, &s->mat, rng, map, score);
	if (*event == ORC_EV_NONE)
		return ORC_ALIVE;

	s->last_tri = -1;
	const orc_material* m = &w->materials[s->mat];
	if (*event == ORC_EV_ELASTIC) {
		orc_elastic_execute(m, &w->opts, &s->p, rng);
		return ORC_ALIVE;
	}
	orc_particle sp;
	const int made = (m->inel_model == ORC_INEL_PENN)
		? orc_penn_execute(m, &w->opts, &s->p, rng, &sp)
		: orc_kieft_execute(m, &w->opts, &s->p, rng, &sp);
	if (made) {
		se->p = sp;
		se->mat = s->mat;
		se->last_tri = -1;
		se->tag = s->tag;
		se->step = 0;
		se->sid = orc_se_sid(s->sid, s->step);
		*se_made = 1;
	}
	return ORC_ALIVE;
}

int orc_step(const orc_world* w, int semantics, orc_state* s, orc_rng* rng,
	orc_state* se, int* se_made, int* event)
{
	return orc_step_scored(w, semantics, s, rng, se, se_made, event, NULL, NULL);
}

/* --------------------------------------------------------------- drivers -- */
typedef struct { orc_state s; int alive; } ref_slot;

int64_t orc_run_reference_deposits(const orc_world* w, const orc_particle* primaries,
	size_t n, uint32_t seed, orc_detected* out, size_t cap, orc_deposit* dep, size_t dep_cap,
	uint64_t* n_dep, uint64_t* n_steps)
{
	uint64_t nd = 0;
	orc_rng rng;
	orc_rng_init_mt(&rng, seed);
	size_t vcap = 1024, vn = 0;
	ref_slot* v = (ref_slot*)malloc(sizeof(ref_slot) * vcap);
	int* status = (int*)malloc(sizeof(int) * vcap);
	int64_t nout = 0;
	uint64_t steps = 0;
	if (!v || !status) return -2;
	for (size_t i = 0; i < n; ++i) {
		/* cpu_particle_manager::push (cpu_particle_manager.inl:23-45) */
		vn = 0;
		memset(&v[0], 0, sizeof(ref_slot));
		v[0].s.p = primaries[i];
		v[0].s.mat = ORC_VACUUM;
		v[0].s.last_tri = -1;
		v[0].s.tag = (uint32_t)i;
		v[0].s.sid = i;
		status[0] = ORC_ALIVE;
		vn = 1;
		/* simple_cpu_driver::simulate_to_end (simple_cpu_driver.h:53-65) */
		for (size_t idx = 0; idx < vn; ++idx) {
			while (status[idx] == ORC_ALIVE) {
				orc_state se;
				int made, ev;
				++steps;
				const float e0 = v[idx].s.p.E;
				status[idx] = orc_step(w, 0, &v[idx].s, &rng, &se, &made, &ev);
				/* energydeposit_cpu_driver.h:79-89: scattering events only */
				if (dep && (ev == ORC_EV_ELASTIC || ev == ORC_EV_INELASTIC) && v[idx].s.p.E != e0) {
					if (nd < dep_cap) {
						dep[nd].pos = v[idx].s.p.pos;
						dep[nd].energy = e0;
						dep[nd].deposit = e0 - v[idx].s.p.E;
						dep[nd].tag = v[idx].s.tag;
						dep[nd].sid = v[idx].s.sid;
					}
					++nd;
				}
				if (made) {
					if (vn == vcap) {
						vcap *= 2;
						v = (ref_slot*)realloc(v, sizeof(ref_slot) * vcap);
						status = (int*)realloc(status, sizeof(int) * vcap);
						if (!v || !status) return -2;
					}
					v[vn].s = se;
					status[vn] = ORC_ALIVE;
					++vn;
				}
			}
		}
		/* flush_detected in vector order (cpu_particle_manager.inl:47-60) */
		for (size_t idx = 0; idx < vn; ++idx) {
			if (status[idx] != ORC_DETECTED) continue;
			if ((size_t)nout >= cap) { free(v); free(status); return -1; }
			out[nout].p = v[idx].s.p;
			out[nout].tag = v[idx].s.tag;
			out[nout].sid = v[idx].s.sid;
			++nout;
		}
	}
	free(v);
	free(status);
	if (n_steps) *n_steps = steps;
	if (n_dep) *n_dep = nd;
	if (nd > dep_cap) return -1;
	return nout;
}

int64_t orc_run_reference(const orc_world* w, const orc_particle* primaries,
	size_t n, uint32_t seed, orc_detected* out, size_t cap, uint64_t* n_steps)
{
	return orc_run_reference_deposits(w, primaries, n, seed, out, cap, NULL, 0, NULL, n_steps);
}

typedef struct {
	const orc_world* w;
	const orc_particle* primaries;
	size_t n;
	uint64_t seed, sid_base;
	uint32_t tag_base;
	const orc_score_map* map;     /* NULL: no scores */
	atomic_size_t next;
	orc_detected* out;
	double* out_scores;
	orc_deposit* dep;             /* NULL: no deposit records */
	size_t dep_cap;
	atomic_ullong ndep;
	size_t cap;
	atomic_llong nout;
	atomic_ullong steps;
	atomic_int overflow;
} philox_job;

/* One entry of a worker's depth-first stack: a particle and its score. */
typedef struct { orc_state s; double score[ORC_MAX_PARAMS]; } scored_state;

static void* philox_worker(void* arg)
{
	philox_job* job = (philox_job*)arg;
	const orc_score_map* map = job->map;
	const int np = map ? map->n_params : 0;
	orc_rng rng;
	orc_rng_init_philox(&rng, job->seed);
	size_t scap = 256;
	scored_state* stack = (scored_state*)malloc(sizeof(scored_state) * scap);
	uint64_t steps = 0;
	for (;;) {
		const size_t i = atomic_fetch_add(&job->next, 64);
		if (i >= job->n) break;
		const size_t iend = i + 64 < job->n ? i + 64 : job->n;
		for (size_t k = i; k < iend; ++k) {
			size_t sn = 1;
			memset(&stack[0], 0, sizeof(scored_state));
			stack[0].s.p = job->primaries[k];
			stack[0].s.mat = ORC_VACUUM;
			stack[0].s.last_tri = -1;
			stack[0].s.tag = job->tag_base + (uint32_t)k;
			stack[0].s.sid = job->sid_base + k;
			while (sn > 0) {
				scored_state cur = stack[--sn];
				orc_state s = cur.s;
				for (;;) {
					orc_state se;
					int made, ev;
					orc_rng_set_stream(&rng, s.sid, s.step);
					++steps;
					const float e0 = s.p.E;
					const int st = orc_step_scored(job->w, 1, &s, &rng, &se, &made, &ev, map, cur.score);
					if (job->dep && (ev == ORC_EV_ELASTIC || ev == ORC_EV_INELASTIC) && s.p.E != e0) {
						const unsigned long long slot = atomic_fetch_add(&job->ndep, 1);
						if (slot < job->dep_cap) {
							orc_deposit* r = &job->dep[slot];
							r->pos = s.p.pos;
							r->energy = e0;
							r->deposit = e0 - s.p.E;
							r->tag = s.tag;
							r->sid = s.sid;
						}
					}
					s.step++;
					if (made) {
						if (sn == scap) {
							scap *= 2;
							stack = (scored_state*)realloc(stack, sizeof(scored_state) * scap);
						}
						/* a secondary inherits its parent's score so far */
						stack[sn].s = se;
						memcpy(stack[sn].score, cur.score, sizeof(cur.score));
						++sn;
					}
					if (st == ORC_ALIVE) continue;
					if (st == ORC_DETECTED) {
						const long long slot = atomic_fetch_add(&job->nout, 1);
						if ((size_t)slot < job->cap) {
							job->out[slot].p = s.p;
							job->out[slot].tag = s.tag;
							job->out[slot].sid = s.sid;
							for (int q = 0; q < np; ++q)
								job->out_scores[(size_t)slot * (size_t)np + (size_t)q] = cur.score[q];
						} else {
							atomic_store(&job->overflow, 1);
						}
					}
					break;
				}
			}
		}
	}
	free(stack);
	atomic_fetch_add(&job->steps, steps);
	return NULL;
}

int64_t orc_run_philox_scored(const orc_world* w, const orc_particle* primaries,
	size_t n, uint64_t seed, uint64_t sid_base, uint32_t tag_base, int threads,
	const orc_score_map* map, orc_detected* out, double* out_scores, size_t cap,
	uint64_t* n_steps)
{
	if (map && (map->n_params < 0 || map->n_params > ORC_MAX_PARAMS))
		return -2;
	philox_job job;
	memset(&job, 0, sizeof(job));
	job.w = w; job.primaries = primaries; job.n = n;
	job.seed = seed; job.sid_base = sid_base; job.tag_base = tag_base;
	job.map = map; job.out = out; job.out_scores = out_scores; job.cap = cap;
	atomic_init(&job.next, 0);
	atomic_init(&job.nout, 0);
	atomic_init(&job.steps, 0);
	atomic_init(&job.overflow, 0);
	if (threads < 1) threads = 1;
	pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
	for (int t = 0; t < threads; ++t)
		pthread_create(&th[t], NULL, philox_worker, &job);
	for (int t = 0; t < threads; ++t)
		pthread_join(th[t], NULL);
	free(th);
	if (n_steps) *n_steps = atomic_load(&job.steps);
	if (atomic_load(&job.overflow)) return -1;
	return (int64_t)atomic_load(&job.nout);
}

int64_t orc_run_philox_deposits(const orc_world* w, const orc_particle* primaries,
	size_t n, uint64_t seed, uint64_t sid_base, uint32_t tag_base, int threads,
	orc_detected* out, size_t cap, orc_deposit* dep, size_t dep_cap, uint64_t* n_dep,
	uint64_t* n_steps)
{
	philox_job job;
	memset(&job, 0, sizeof(job));
	job.w = w; job.primaries = primaries; job.n = n;
	job.seed = seed; job.sid_base = sid_base; job.tag_base = tag_base;
	job.out = out; job.cap = cap; job.dep = dep; job.dep_cap = dep_cap;
	atomic_init(&job.next, 0);
	atomic_init(&job.nout, 0);
	atomic_init(&job.ndep, 0);
	atomic_init(&job.steps, 0);
	atomic_init(&job.overflow, 0);
	if (threads < 1) threads = 1;
	pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
	for (int t = 0; t < threads; ++t)
		pthread_create(&th[t], NULL, philox_worker, &job);
	for (int t = 0; t < threads; ++t)
		pthread_join(th[t], NULL);
	free(th);
	if (n_steps) *n_steps = atomic_load(&job.steps);
	if (n_dep) *n_dep = atomic_load(&job.ndep);
	if (atomic_load(&job.overflow) || atomic_load(&job.ndep) > dep_cap) return -1;
	return (int64_t)atomic_load(&job.nout);
}

int64_t orc_run_philox(const orc_world* w, const orc_particle* primaries,
	size_t n, uint64_t seed, uint64_t sid_base, uint32_t tag_base, int threads,
	orc_detected* out, size_t cap, uint64_t* n_steps)
{
	return orc_run_philox_scored(w, primaries, n, seed, sid_base, tag_base, threads, NULL, out, NULL,
		cap, n_steps);
}

void orc_pf_eval_batch(int fn, const float* a, const float* b, float* out, size_t n)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = orc_pf_eval(fn, a[i], b ? b[i] : 0.0f);
}
