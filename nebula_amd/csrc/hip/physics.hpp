// Device-side tracking physics for gfx950.
//
// Every floating-point expression keeps the operation order of the reference
// function cited beside it (paths relative to the reference's source/), and
// the library is compiled with -ffp-contract=off, so each step is
// bit-reproducible against a CPU evaluation of the same expressions.
// Transcendentals are the float routines of this file (see "math").
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace nbl { namespace dev {

constexpr float EPS = 10.0f * 1.1920928955078125e-07f;  // config/data_types.h:55
constexpr float PI_F = 3.1415926535897932f;             // config/config.h:32
// Bounds on the two unbounded loops of the reference (octree cell walk,
// plasmon-decay rejection), so that every wave finishes.
constexpr uint32_t MAX_CELLS = 1000000u;
constexpr int MAX_REJECT = 1000;

// Special material indices (core/material_manager.h:77-86).
constexpr int MAT_NOP = -128, MAT_TERMINATOR = -127, MAT_DETECTOR = -126,
	MAT_DETECTOR_LT50 = -125, MAT_DETECTOR_GE50 = -124, MAT_VACUUM = -123,
	MAT_MIRROR = -122;

// Physics option bits (same values as NBL_PHYS_* in nebula.h).
constexpr uint32_t PH_EL_PHONON = 1u << 0, PH_EL_RECOIL = 1u << 1, PH_PN_SE = 1u << 2,
	PH_KI_OPTICAL = 1u << 3, PH_KI_SE = 1u << 4, PH_KI_INST = 1u << 5, PH_KI_MOMCONS = 1u << 6,
	PH_BI_QUANTUM = 1u << 7, PH_BI_REFRACT = 1u << 8, PH_BI_ABSORB = 1u << 9;

enum { ALIVE = 0, TERMINATED = 1, DETECTED = 2 };
enum { EV_NONE = 0, EV_INELASTIC = 1, EV_ELASTIC = 2, EV_INTERSECT = 3 };
// Particle flags.
constexpr uint32_t FL_PEND_INEL = 1u, FL_PEND_DET = 2u;
// Bits 16..31 of a particle's flags: its world (geometry + material set) in
// a driver that tracks several scans at once.
constexpr uint32_t FL_WORLD_SHIFT = 16u, FL_WORLD_MASK = 0xffff0000u;

struct F3 { float x, y, z; };

// ------------------------------------------------------------------ math --
// Float transcendentals built from IEEE float +, -, *, / and integer bit
// operations only, so that (compiled without FMA contraction) they give the
// same bits on the GPU and in a CPU evaluation.  Methods: exponent/mantissa
// split with log(1+f) = f - f^2/2 + s(f^2/2 + R(s^2)), s = f/(2+f), R the
// atanh series; Cody-Waite reduction by k ln2 or k pi/2 with split constants
// and Taylor polynomials; arctangent by reduction to |u| <= tan(pi/8) and a
// least-squares polynomial; cube root by a bit-level first guess and four
// Newton steps.  All within 2.5 ulp of the exact functions.
__device__ __forceinline__ float pf_nan() { return __uint_as_float(0x7fc00000u); }
__device__ __forceinline__ float pf_pow2(int k) { return __uint_as_float((uint32_t)(k + 127) << 23); }
constexpr float PF_P1 = 0x1.922p+0f, PF_P2 = -0x1.2aep-18f, PF_P3 = -0x1.de973ep-31f;
constexpr float PF_LN2_HI = 0x1.62e4p-1f, PF_LN2_LO = 0x1.7f7d1cp-20f, PF_INV_LN2 = 0x1.715476p+0f;
constexpr float PF_PI_HI = 0x1.921fb6p+1f, PF_PI_LO = -0x1.777a5cp-24f;
constexpr float PF_PI_2_HI = 0x1.921fb6p+0f, PF_PI_2_LO = -0x1.777a5cp-25f;
constexpr float PF_PI_4_HI = 0x1.921fb6p-1f, PF_PI_4_LO = -0x1.777a5cp-26f;
constexpr float PF_TWO_OVER_PI = 0x1.45f306p-1f, PF_TAN_PI_8 = 0x1.a8279ap-2f;

__device__ inline float pf_log(float x)
{
	if (x != x) return x;
	if (x < 0.0f) return pf_nan();
	if (x == 0.0f) return -__builtin_inff();
	uint32_t i = __float_as_uint(x);
	if (i >= 0x7f800000u) return x;
	int e = 0;
	if (i < 0x00800000u) { x = x * 0x1p25f; i = __float_as_uint(x); e = -25; }
	e += (int)(i >> 23) - 127;
	const uint32_t m = i & 0x007fffffu;
	if (m > 0x3504f3u) { i = m | 0x3f000000u; e += 1; }
	else i = m | 0x3f800000u;
	const float f = __uint_as_float(i) - 1.0f;
	const float s = f / (2.0f + f);
	const float z = s * s;
	const float R = z * (0x1.555556p-1f + z * (0x1.99999ap-2f + z * (0x1.24924ap-2f + z * 0x1.c71c72p-3f)));
	const float hfsq = 0.5f * f * f;
	const float l1p = f - (hfsq - s * (hfsq + R));
	const float fe = (float)e;
	return fe * PF_LN2_HI + (l1p + fe * PF_LN2_LO);
}

__device__ __forceinline__ float pf_exp_p(float r)
{
	return 0.5f + r * (0x1.555556p-3f + r * (0x1.555556p-5f + r * (0x1.111112p-7f
		+ r * (0x1.6c16c2p-10f + r * 0x1.a01a02p-13f))));
}

__device__ inline float pf_exp(float x)
{
	if (x != x) return x;
	if (x > 0x1.62e42ep+6f) return __builtin_inff();
	if (x < -0x1.9fe368p+6f) return 0.0f;
	const float kf = rintf(x * PF_INV_LN2);
	const int k = (int)kf;
	const float r = (x - kf * PF_LN2_HI) - kf * PF_LN2_LO;
	const float e = 1.0f + (r + (r * r) * pf_exp_p(r));
	if (k > 127) return (e * 0x1p127f) * pf_pow2(k - 127);
	if (k < -126) return (e * 0x1p-126f) * pf_pow2(k + 126);
	return e * pf_pow2(k);
}

__device__ inline float pf_expm1(float x)
{
	if (x != x) return x;
	if (x > 0x1.62e42ep+6f) return __builtin_inff();
	if (x < -0x1.154246p+4f) return -1.0f;
	if (fabsf(x) < 0x1.62e430p-2f) {
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

__device__ inline float pf_cbrt(float x)
{
	if (x != x || x == 0.0f || fabsf(x) == __builtin_inff()) return x;
	float ax = fabsf(x);
	float post = 1.0f;
	if (ax < 0x1p-126f) { ax = ax * 0x1p24f; post = 0x1p-8f; }
	float y = __uint_as_float(__float_as_uint(ax) / 3u + 0x2a555555u);
#pragma unroll
	for (int it = 0; it < 4; ++it)
		y = (2.0f * y + ax / (y * y)) / 3.0f;
	y = y * post;
	return copysignf(y, x);
}

__device__ __forceinline__ float pf_sin_r(float r)
{
	const float z = r * r;
	return r + (r * z) * (-0x1.555556p-3f + z * (0x1.111112p-7f + z * (-0x1.a01a02p-13f
		+ z * (0x1.71de3ap-19f + z * -0x1.ae6456p-26f))));
}
__device__ __forceinline__ float pf_cos_r(float r)
{
	const float z = r * r;
	return 1.0f - (0.5f * z - (z * z) * (0x1.555556p-5f + z * (-0x1.6c16c2p-10f
		+ z * (0x1.a01a02p-16f + z * -0x1.27e4fcp-22f))));
}
__device__ __forceinline__ float pf_reduce(float x, int& q)
{
	const float kf = rintf(x * PF_TWO_OVER_PI);
	q = ((int)kf) & 3;
	return ((x - kf * PF_P1) - kf * PF_P2) - kf * PF_P3;
}
// sin and cos of the same argument (the reference calls sincos on the CPU
// as separate sin and cos, which give the same values as these)
__device__ inline void pf_sincos(float x, float& s, float& c)
{
	if (x != x || fabsf(x) == __builtin_inff()) { s = pf_nan(); c = pf_nan(); return; }
	int q;
	const float r = pf_reduce(x, q);
	const float sr = pf_sin_r(r), cr = pf_cos_r(r);
	switch (q) {
	case 0: s = sr; c = cr; break;
	case 1: s = cr; c = -sr; break;
	case 2: s = -sr; c = -cr; break;
	default: s = -cr; c = sr; break;
	}
}

__device__ __forceinline__ float pf_atan_r(float u)
{
	const float z = u * u;
	return u + (u * z) * (-0x1.555556p-2f + z * (0x1.999982p-3f + z * (-0x1.248b8p-3f
		+ z * (0x1.c5adfep-4f + z * (-0x1.62ae6ep-4f + z * 0x1.a49afap-5f)))));
}
__device__ inline float pf_atan2(float y, float x)
{
	if (x != x || y != y) return x + y;
	const float ax = fabsf(x), ay = fabsf(y);
	const bool xneg = (__float_as_uint(x) >> 31) != 0;
	if (ay == 0.0f)
		return xneg ? copysignf(PF_PI_HI, y) : y;
	if (ax == 0.0f)
		return copysignf(PF_PI_2_HI, y);
	const float inf = __builtin_inff();
	if (ax == inf || ay == inf) {
		float a;
		if (ax == inf && ay == inf) a = xneg ? 3.0f * PF_PI_4_HI : PF_PI_4_HI;
		else if (ax == inf) a = xneg ? PF_PI_HI : 0.0f;
		else a = PF_PI_2_HI;
		return copysignf(a, y);
	}
	const bool swap = ay > ax;
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

__device__ __forceinline__ float m_log(float x) { return pf_log(x); }
__device__ __forceinline__ float m_exp(float x) { return pf_exp(x); }
__device__ __forceinline__ float m_expm1(float x) { return pf_expm1(x); }
__device__ __forceinline__ float m_cbrt(float x) { return pf_cbrt(x); }
__device__ __forceinline__ float m_atan2(float y, float x) { return pf_atan2(y, x); }

// common/scalar_math.h:16-25; float->int conversions follow the x86 CPU
// build for out-of-range and NaN inputs (INT_MIN).
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
__device__ __forceinline__ float clampr(float i, float lo, float hi) { return i < lo ? lo : (i > hi ? hi : i); }
__device__ __forceinline__ float fmin_std(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_std(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ int to_int_x86(float f)
{
	return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : (int)0x80000000;
}
__device__ __forceinline__ int floorr(float x) { return to_int_x86(floorf(x)); }
__device__ __forceinline__ int rintr(float x) { return to_int_x86(rintf(x)); }
__device__ __forceinline__ float saturater(float x) { return clampr(x, 0.0f, 1.0f); }

// common/vector_math.h
__device__ __forceinline__ F3 v_add(F3 a, F3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
__device__ __forceinline__ F3 v_sub(F3 a, F3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ F3 v_mul(F3 a, float b) { return { a.x * b, a.y * b, a.z * b }; }
__device__ __forceinline__ F3 s_mul(float a, F3 b) { return { a * b.x, a * b.y, a * b.z }; }
__device__ __forceinline__ F3 cross(F3 a, F3 b)
{
	return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float mag2(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// normalised(): a * (1/sqrt(a.a)) with IEEE sqrt and division.
__device__ __forceinline__ F3 normalised(F3 a)
{
	const float r = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
	return v_mul(a, r);
}
// vector_math.h:171-186
__device__ inline F3 make_normal_vec(F3 dir, float phi)
{
	// cos and sin of the azimuth atan2(y, x) as x/r and y/r (equal up to
	// rounding; r = 0 takes the atan2 limits: cosine +-1 by the sign of x)
	const float r = sqrtf(dir.x * dir.x + dir.y * dir.y);
	float sin_az = 0.0f, cos_az = signbit(dir.x) ? -1.0f : 1.0f;
	if (r > 0.0f) {
		cos_az = dir.x / r;
		sin_az = dir.y / r;
	}
	const F3 unit_v = { dir.z * cos_az, dir.z * sin_az, -sqrtf(dir.x * dir.x + dir.y * dir.y) };
	const F3 unit_u = cross(unit_v, dir);
	float sin_phi, cos_phi;
	pf_sincos(phi, sin_phi, cos_phi);
	return v_add(v_mul(unit_u, cos_phi), v_mul(unit_v, sin_phi));
}
// vector_math.h:194-206
__device__ inline F3 make_unit_vec(float cos_theta, float phi)
{
	float sin_phi, cos_phi;
	pf_sincos(phi, sin_phi, cos_phi);
	cos_theta = clampr(cos_theta, -1.0f, 1.0f);
	const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
	return { sin_theta * cos_phi, sin_theta * sin_phi, cos_theta };
}

// ------------------------------------------------------------------- RNG --
// Counter-based Philox4x32-10 (Salmon et al., SC'11): key = run seed,
// counter = (sid lo, sid hi, step, block of four draws).  A particle's draws
// depend only on (seed, sid, step, draw index), so trajectories are
// independent of scheduling and can be replayed.
struct Rng {
	uint32_t k0, k1, c0, c1, c2, idx;
	uint32_t b0, b1, b2, b3;
};

__device__ __forceinline__ void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
	uint32_t k0, uint32_t k1)
{
#pragma unroll
	for (int r = 0; r < 10; ++r) {
		const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
		const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
		const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
		c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
}

__device__ __forceinline__ void rng_start(Rng& r, uint32_t k0, uint32_t k1, uint64_t sid, uint32_t step)
{
	r.k0 = k0; r.k1 = k1;
	r.c0 = (uint32_t)sid; r.c1 = (uint32_t)(sid >> 32); r.c2 = step;
	r.idx = 0;
}

// Uniform in (0, 1]: ((w >> 8) + 1) * 2^-24, exact in float.
__device__ __forceinline__ float rng_unit(Rng& r)
{
	const uint32_t i = r.idx & 3u;
	if (i == 0) {
		uint32_t a = r.c0, b = r.c1, c = r.c2, d = r.idx >> 2;
		philox4x32_10(a, b, c, d, r.k0, r.k1);
		r.b0 = a; r.b1 = b; r.b2 = c; r.b3 = d;
	}
	const uint32_t w = (i == 0) ? r.b0 : (i == 1) ? r.b1 : (i == 2) ? r.b2 : r.b3;
	r.idx++;
	return (float)((w >> 8) + 1u) * (1.0f / 16777216.0f);
}

// Resume a stream at draw `idx` (re-running an event whose path draws were
// taken by an earlier launch).
__device__ __forceinline__ void rng_seek(Rng& r, uint32_t idx)
{
	const uint32_t blk = idx >> 2;
	uint32_t a = r.c0, b = r.c1, c = r.c2, d = blk;
	philox4x32_10(a, b, c, d, r.k0, r.k1);
	r.b0 = a; r.b1 = b; r.b2 = c; r.b3 = d;
	r.idx = idx;   // rng_unit recomputes the block if idx is a multiple of four
}

// common/util/random.h:56-84: phi = 2 pi U; uniform_vector draws phi first.
__device__ __forceinline__ float rng_phi(Rng& r) { return (2.0f * PI_F) * rng_unit(r); }
__device__ __forceinline__ F3 rng_uniform_vector(Rng& r)
{
	const float phi = rng_phi(r);
	const float ct = 2.0f * rng_unit(r) - 1.0f;
	return make_unit_vec(ct, phi);
}

// Stream id of a secondary electron: splitmix64 finaliser of the parent's
// (sid, step), top bit set.
__device__ __forceinline__ uint64_t se_sid(uint64_t parent_sid, uint32_t parent_step)
{
	uint64_t z = parent_sid * 0x9E3779B97F4A7C15ull + (uint64_t)parent_step + 1u;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	z ^= z >> 31;
	return z | 0x8000000000000000ull;
}

// ---------------------------------------------------------------- tables --
struct Table1 { const float* d; int n; float xmin, xstep; };
struct Table2 { const float* d; int nx, ny; float xmin, xstep, ymin, ystep; };
struct Table3 { const float* d; int nx, ny, nz; float xmin, xstep, ymin, ystep, zmin, zstep; };

// common/util/table_1D.inl:66-81
__device__ __forceinline__ float t1_get(const Table1& t, float x)
{
	const float xi = (x - t.xmin) * t.xstep;
	const int low = clampi(floorr(xi), 0, t.n - 2);
	const float lv = t.d[low], hv = t.d[low + 1];
	const float frac = xi - (float)low;
	return (1.0f - frac) * lv + frac * hv;
}

// common/util/table_2D.inl:73-94 (row-major [x][y])
__device__ __forceinline__ float t2_get(const Table2& t, float x, float y)
{
	const float xi = (x - t.xmin) * t.xstep;
	const float yi = (y - t.ymin) * t.ystep;
	const int lx = clampi(floorr(xi), 0, t.nx - 2);
	const int ly = clampi(floorr(yi), 0, t.ny - 2);
	const float* r0 = t.d + (size_t)lx * t.ny + ly;
	const float* r1 = r0 + t.ny;
	const float v00 = r0[0], v10 = r1[0], v01 = r0[1], v11 = r1[1];
	const float fx = xi - (float)lx, fy = yi - (float)ly;
	return (1.0f - fx) * (1.0f - fy) * v00
		+ fx * (1.0f - fy) * v10
		+ (1.0f - fx) * fy * v01
		+ fx * fy * v11;
}

// common/util/table_2D.inl:96-114
__device__ __forceinline__ float t2_get_rounddown(const Table2& t, float x, float y)
{
	const float xi = (x - t.xmin) * t.xstep;
	const float yi = (y - t.ymin) * t.ystep;
	const int lx = clampi(floorr(xi), 0, t.nx - 1);
	const int ly = clampi(floorr(yi), 0, t.ny - 1);
	return t.d[(size_t)lx * t.ny + ly];
}
__device__ __forceinline__ float t2_get_nearest(const Table2& t, float x, float y)
{
	const float xi = (x - t.xmin) * t.xstep;
	const float yi = (y - t.ymin) * t.ystep;
	const int nx = clampi(rintr(xi), 0, t.nx - 1);
	const int ny = clampi(rintr(yi), 0, t.ny - 1);
	return t.d[(size_t)nx * t.ny + ny];
}

// common/util/table_3D.inl:78-121 ([x][y][z])
__device__ __forceinline__ float t3_get(const Table3& t, float x, float y, float z)
{
	const float xi = (x - t.xmin) * t.xstep;
	const float yi = (y - t.ymin) * t.ystep;
	const float zi = (z - t.zmin) * t.zstep;
	const int lx = clampi(floorr(xi), 0, t.nx - 2);
	const int ly = clampi(floorr(yi), 0, t.ny - 2);
	const int lz = clampi(floorr(zi), 0, t.nz - 2);
	const size_t sy = (size_t)t.nz, sx = (size_t)t.ny * t.nz;
	const float* p = t.d + (size_t)lx * sx + (size_t)ly * sy + lz;
	const float v000 = p[0], v001 = p[1];
	const float v010 = p[sy], v011 = p[sy + 1];
	const float v100 = p[sx], v101 = p[sx + 1];
	const float v110 = p[sx + sy], v111 = p[sx + sy + 1];
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
__device__ __forceinline__ float t3_get_rounddown(const Table3& t, float x, float y, float z)
{
	const float xi = (x - t.xmin) * t.xstep;
	const float yi = (y - t.ymin) * t.ystep;
	const float zi = (z - t.zmin) * t.zstep;
	const int lx = clampi(floorr(xi), 0, t.nx - 1);
	const int ly = clampi(floorr(yi), 0, t.ny - 1);
	const int lz = clampi(floorr(zi), 0, t.nz - 1);
	return t.d[((size_t)lx * t.ny + ly) * t.nz + lz];
}

// -------------------------------------------------------------- material --
struct Material {
	float barrier, phonon_loss, recoil_const, fermi, band_gap;
	// mean-free-path scales of the derivative estimator (1 = the tables)
	float el_scale, in_scale;
	int inel_model;
	Table1 el_imfp, in_imfp;
	Table2 el_icdf, in_icdf, ki_ion;
	Table3 pn_q, pn_ion;
};

// -------------------------------------------------------------- geometry --
// Triangle as three float4: (r0.xyz, e1.x), (e1.yz, e2.xy), (e2.z, mat_in,
// mat_out, 0) with the materials stored as int bits.  (A 64-byte record that
// also holds e1 x e2 measured slower: the extra load costs more than the
// cross product it saves.)
struct Tri { F3 r0, e1, e2; int mat_in, mat_out; };

__device__ __forceinline__ Tri load_tri(const float4* t, int i)
{
	const float4 a = t[3 * i], b = t[3 * i + 1], c = t[3 * i + 2];
	Tri r;
	r.r0 = { a.x, a.y, a.z };
	r.e1 = { a.w, b.x, b.y };
	r.e2 = { b.z, b.w, c.x };
	r.mat_in = __float_as_int(c.y);
	r.mat_out = __float_as_int(c.z);
	return r;
}

struct Geometry {
	const int* od;
	const float4* tris;
	// Screening record per entry of the octree table, two float4, filled at
	// every position that names a triangle: (r0, |e1|_1 |e2|_1) and (e1 x e2
	// as the triangle test computes it, material_in | material_out as two
	// int16 in the bits of w).  Stored in list order, so the screen reads
	// neighbouring records for neighbouring pairs, without waiting for the
	// triangle index.
	const float4* scr;
	// odc[i] = (od[i], length of the leaf list at -od[i] when od[i] < 0
	// names one): the walk reads a child entry and its list length in one
	// load.
	const int2* odc;
	F3 center, halfsize;
	float max_extent;
	int ntri;
	// Bound on the rounding of one move inside the domain (a position
	// update's error, summed over components, plus the safety's own
	// subtraction): 3e-7 x the domain's largest 1-norm + 1e-6 nm.
	float step_eps;
};

// core/triangle.inl:7-28
__device__ __forceinline__ float intersect_ray(const Tri& t, F3 start, F3 dir)
{
	const F3 pvec = cross(dir, t.e2);
	const float det = dot(t.e1, pvec);
	if (fabsf(det) < EPS)
		return -1.0f;
	const F3 s = v_sub(start, t.r0);
	const float u = dot(s, pvec) / det;
	if ((u < -EPS) || (u > 1.0f + EPS))
		return -1.0f;
	const F3 qvec = cross(s, t.e1);
	const float v = dot(dir, qvec) / det;
	if ((v < -EPS) || (u + v > 1.0f + EPS))
		return -1.0f;
	const float tt = dot(t.e2, qvec) / det;
	if (tt < 0.0f)
		return -1.0f;
	return tt;
}

// geometry/octree.inl:24-30
__device__ __forceinline__ bool in_domain(const Geometry& g, F3 p)
{
	const F3 c = g.center, h = g.halfsize;
	return (p.x > c.x - h.x) && (p.x < c.x + h.x)
		&& (p.y > c.y - h.y) && (p.y < c.y + h.y)
		&& (p.z > c.z - h.z) && (p.z < c.z + h.z);
}

// Levels of the octree table cached per thread: path[(l - 1) * stride] holds
// the node index at level l (1..21) of the current cell.
constexpr int PATH_LEVELS = 21;

// The pieces of the cell walk of geometry/octree.inl:32-192.
//
// The reference re-descends from the root for every cell.  Here the node
// indices of the current cell's ancestors are kept in LDS (`path`, one int
// per level, lane-strided), so moving to a neighbour loads only the flipped
// level and whatever lies below it.  The cell-centre arithmetic still runs
// from the root in the reference's order, so every value is unchanged.
//
// cell_find: centre (ax, ay, az) and scale aw of the leaf holding `cur`,
// extending `location` down to it; returns the start of its triangle list
// and its length in `count`.
__device__ __forceinline__ int cell_find(const Geometry& g, F3 cur, uint64_t& location,
	float& ax, float& ay, float& az, float& aw, int* __restrict__ path, int stride, int& count)
{
	const int2* __restrict__ odc = g.odc;
	const F3 hs = g.halfsize;
	// The reference adds aw * hs * (bit - 0.5) per level.  The product with
	// +-0.5 is exact, and scaling by a power of two commutes with rounding
	// (aw >= 2^-21 and hs >= 1 nm keep every product far from underflow), so
	// it equals +-(hw * hs) with hw = aw / 2: the same bits in three
	// operations per axis instead of six.
	ax = g.center.x; ay = g.center.y; az = g.center.z;
	float hw = 0.5f;
	const int clz = __clzll((long long)location);
	for (int i = 60 - clz; i >= 0; i -= 3) {
		const unsigned oct = (unsigned)(location >> i) & 7u;
		const float hx = hw * hs.x, hy = hw * hs.y, hz = hw * hs.z;
		ax += (oct & 1u) ? hx : -hx;
		ay += (oct & 2u) ? hy : -hy;
		az += (oct & 4u) ? hz : -hz;
		hw *= 0.5f;
	}
	// levels 1..L-1 are unchanged ancestors; level L holds the new octant
	int lev = (63 - clz) / 3;
	int index = 0;
	count = 0;
	if (lev > 0) {
		const int parent = (lev >= 2) ? path[(lev - 2) * stride] : 0;
		const int2 e = odc[parent + (int)((unsigned)location & 7u)];
		index = e.x;
		count = e.y;
		path[(lev - 1) * stride] = index;
	}
	while (index >= 0) {
		unsigned oct = 0;
		oct ^= (cur.x > ax) ? 1u : 0u;
		oct ^= (cur.y > ay) ? 2u : 0u;
		oct ^= (cur.z > az) ? 4u : 0u;
		location = (location << 3) | oct;
		const int2 e = odc[index + (int)oct];
		index = e.x;
		count = e.y;
		path[lev * stride] = index;   // level lev + 1
		++lev;
		const float hx = hw * hs.x, hy = hw * hs.y, hz = hw * hs.z;
		ax += (oct & 1u) ? hx : -hx;
		ay += (oct & 2u) ? hy : -hy;
		az += (oct & 4u) ? hz : -hz;
		hw *= 0.5f;
	}
	aw = 2.0f * hw;
	return -index;
}

// Distance to the cell wall the ray leaves through (target -1/-2/-4 for the
// x/y/z wall).
__device__ __forceinline__ void cell_exit(const Geometry& g, F3 cur, F3 dir, float ax, float ay, float az,
	float aw, float& isect, int& target)
{
	const F3 hs = g.halfsize;
	const float hx = aw * hs.x, hy = aw * hs.y, hz = aw * hs.z;
	const float tx = (ax + copysignf(hx + EPS, dir.x) - cur.x) / dir.x;
	const float ty = (ay + copysignf(hy + EPS, dir.y) - cur.y) / dir.y;
	const float tz = (az + copysignf(hz + EPS, dir.z) - cur.z) / dir.z;
	if ((tx < ty) && (tx < tz)) { isect = tx; target = -1; }
	else if ((ty < tx) && (ty < tz)) { isect = ty; target = -2; }
	else { isect = tz; target = -4; }
}

// Candidate test: octree.inl:128-138 skips triangles whose outgoing material
// is the current one, then core/triangle.inl gives the distance (or -1).
// (A separate 16-byte record with e1 x e2 and the materials, read before the
// triangle, measured 9% slower: the extra dependent load costs more than the
// cross product it saves.  Estimating the three quotients with one
// reciprocal and dividing only where the outcome is not certain kept every
// result but measured 3% slower in the wave-wide walk.)
__device__ __forceinline__ float cell_candidate(const Geometry& g, int ti, F3 cur, F3 dir, int ignore_mat)
{
	const Tri tri = load_tri(g.tris, ti);
	const int mat_out = (dot(cross(tri.e1, tri.e2), dir) < 0.0f) ? tri.mat_in : tri.mat_out;
	if (mat_out == ignore_mat)
		return -1.0f;
	return intersect_ray(tri, cur, dir);
}

// First stage of the wave-wide walk's candidate test, on the 32-byte
// screening record of the list entry: the material test of cell_candidate (the same e1 x e2,
// computed on the host with the same operations, so the same result) and a
// plane test that drops
// a pair only when its intersection distance cannot matter.  For the pair's
// float inputs (s = cur - r0, dir, e1, e2), Moller-Trumbore's t is
// t* = -(s . n) / (dir . n) with n = e1 x e2 in exact arithmetic, and
// f(x) = s . n + x (dir . n) is the plane value at distance x along the ray.
// The pair is dropped when f has one sign, by more than
// tol = 1e-3 (|s| + L) |e1||e2|, at both x = -m and x = L with
// m = 1e-2 (1 + |s|) and L = 1.01 to_go + m: then t* < -m or t* > L.
// (Lengths here are 1-norms, which bound the 2-norms from above: every
// margin only grows.)
// Rounding moves f by at most about 1e-5 (|s| + L) |e1||e2| (a few unit
// roundoffs times the terms' magnitudes), far inside tol.  Pairs whose ray
// is within |dir . n| < 1e-2 |e1||e2| of the plane always pass: elsewhere
// the computed determinant and numerator of intersect_ray are within
// 1e-4 |t*| + 1e-4 |s| of their exact quotient, so a dropped pair has
// computed t < 0 (cannot hit) or t > to_go + EPS.  Such a far result cannot
// change the running-minimum fold: taking it needs isect >= t - EPS > to_go,
// and from there the owner ends without a hit whatever it takes next (a
// candidate below to_go is taken in both folds and makes them equal again).
// Pairs that pass are tested in full (intersect_ray) in the second stage.
__device__ __forceinline__ float norm1(F3 a) { return fabsf(a.x) + fabsf(a.y) + fabsf(a.z); }
__device__ __forceinline__ bool cell_may_hit(const Geometry& g, int entry, F3 cur, F3 dir, int ignore_mat,
	float to_go)
{
	const float4 a = g.scr[2 * entry], b = g.scr[2 * entry + 1];
	const F3 nrm = { b.x, b.y, b.z };
	const float dn = dot(nrm, dir);
	const uint32_t mats = __float_as_uint(b.w);
	const int mat_out = (dn < 0.0f) ? (int)(int16_t)(mats & 0xffffu) : (int)(int16_t)(mats >> 16);
	if (mat_out == ignore_mat)
		return false;
	const float ee = a.w;
	if (!(fabsf(dn) >= 1e-2f * ee))
		return true;
	const F3 sv = { cur.x - a.x, cur.y - a.y, cur.z - a.z };
	const float slen = norm1(sv);
	const float d0 = dot(sv, nrm);
	const float m = 1e-2f * (1.0f + slen);
	const float L = 1.01f * to_go + m;
	const float fa = d0 - m * dn, fb = d0 + L * dn;
	const float tol = 1e-3f * (slen + L) * ee;
	return !((fa > tol && fb > tol) || (fa < -tol && fb < -tol));
}

// Safety of a walk's start point: a radius S such that any later ray of
// length d starting within distance r of `cur`, with r + d <= S, finds no
// hit in the walk (the walk returns "no hit" and the full free path), so
// that such a step can skip the walk with the same result.
//
// Triangle part, for one triangle of the start cell from its screening
// record: with s = cur - r0, n the exact e1 x e2 and t the computed
// intersect_ray distance of a ray from p' = cur + D (|D| <= r) along any
// unit direction, intersect_ray's numerator and determinant are the exact
// s'.n and -dir.n up to rounding errors of at most 1e-6 |s'|_1 |e1|_1 |e2|_1
// and 1e-6 |e1|_1 |e2|_1 (a few unit roundoffs of each product, with
// |dir|_1 <= sqrt(3)), and the record's n is within 2e-7 |e1|_1 |e2|_1 of
// the exact one.  Hence |t| >= (|s.n| - r |n| - err) / (|n| + err'), and
// |t| >= d holds whenever a - 2e-6 |s|_1 ee >= S (|n|~ (1 + 2e-6) + 3e-6 ee)
// with a = |s.n| as computed here, ee = |e1|_1 |e2|_1 and |n|~ >= |n| the
// record's norm rounded up (the terms in r double the error allowance,
// r <= S).  A computed t >= d (= the remaining free path) is never taken as
// a hit: the walk keeps the nearest t <= (cell exit + EPS) and reports a hit
// only below the free path.  The bound below keeps a further 1e-4 relative
// slack for its own rounding (reciprocal and square root within 1 ulp).
// Triangles skipped by the material or last-triangle tests count too: the
// direction and the last triangle change between steps.
__device__ __forceinline__ float pair_safety(const Geometry& g, int entry, F3 cur)
{
	const float4 a = g.scr[2 * entry], b = g.scr[2 * entry + 1];
	const F3 nrm = { b.x, b.y, b.z };
	const F3 sv = { cur.x - a.x, cur.y - a.y, cur.z - a.z };
	const float ee = a.w;
	const float num = fabsf(dot(sv, nrm)) - 2e-6f * norm1(sv) * ee;
	const float den = sqrtf(dot(nrm, nrm)) * 1.0001f + 1e-5f * ee;
	return num * __builtin_amdgcn_rcpf(den) * 0.9999f;
}

// Wall part: distance from `cur` to the walls of its leaf cell (centre
// a, half extents aw x halfsize), less a margin that covers the rounding of
// cell_exit's exit distances (a few unit roundoffs of the coordinates'
// magnitudes, and EPS pushes the walls outwards): a segment that ends at
// least that far inside the cell is seen to end before the cell's exit.
__device__ __forceinline__ float wall_safety(const Geometry& g, F3 cur, float ax, float ay, float az, float aw)
{
	const F3 hs = g.halfsize;
	const float hx = aw * hs.x, hy = aw * hs.y, hz = aw * hs.z;
	const float w = fminf(fminf(hx - fabsf(cur.x - ax), hy - fabsf(cur.y - ay)), hz - fabsf(cur.z - az));
	const float mag = fabsf(ax) + fabsf(ay) + fabsf(az) + hx + hy + hz;
	return w - (1e-3f + 4e-6f * mag);
}

// The neighbour across wall `target` (-1/-2/-4): flip that axis' bit at the
// deepest level where the move stays inside the parent; location 1 (or 0)
// means the ray left the octree.
__device__ __forceinline__ void cell_next(uint64_t& location, int target, F3 dir)
{
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
}

// geometry/octree.inl:32-192: walk leaf cells by location code; returns the
// hit triangle (or -1) and the distance.  One lane, one ray.
__device__ inline int propagate(const Geometry& g, F3 start, F3 dir, float distance,
	int ignore_tri, int ignore_mat, float& out_dist, int* __restrict__ path, int stride)
{
	const int* __restrict__ od = g.od;
	F3 cur = start;
	float to_go = distance;
	uint64_t location = 1;
	uint32_t cells = 0;
	do {
		float ax, ay, az, aw;
		int count;
		int index = cell_find(g, cur, location, ax, ay, az, aw, path, stride, count);
		++cells;
		float isect;
		int target;
		cell_exit(g, cur, dir, ax, ay, az, aw, isect, target);
		for (;;) {
			const int ti = od[index++];
			if (ti < 0)
				break;
			if (ti == ignore_tri)
				continue;
			const float t = cell_candidate(g, ti, cur, dir, ignore_mat);
			if ((t > 0.0f) && (t <= isect + EPS)) {
				isect = t;
				target = ti;
			}
		}
		if (isect >= to_go) {
			out_dist = distance;
			return -1;
		} else if (target >= 0) {
			out_dist = distance - to_go + isect;
			return target;
		}
		to_go -= isect;
		cur = v_add(cur, v_mul(dir, isect));
		cell_next(location, target, dir);
	} while (location > 1 && cells < MAX_CELLS);
	out_dist = distance;
	return -1;
}

// Lane `src`'s value of a pointer.
__device__ __forceinline__ const void* shfl_ptr(const void* p, int src)
{
	const uint64_t v = (uint64_t)p;
	const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
	const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
	return (const void*)(((uint64_t)hi << 32) | lo);
}

// Sum of v over lanes 0..lane; every lane of the wave must be active.
__device__ __forceinline__ int wave_incl_sum(int v)
{
	const int lane = (int)__lane_id();
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		const int t = __shfl_up(v, d);
		if (lane >= d)
			v += t;
	}
	return v;
}

// Owner of pair base + lane in a round of the wave-wide walk: the last lane
// whose pairs start at or before it (excl = first pair of this lane's list,
// n = its length).  Each lane whose pairs start inside the chunk marks that
// position in `owner_at` (64 ints of LDS); position 0 holds the owner carried
// in from earlier (the last lane with excl <= base).
__device__ __forceinline__ int chunk_owner(int base, int excl, int n, int* owner_at, int lane)
{
	const uint64_t before = __ballot(excl <= base);
	const int o0 = 63 - (int)__builtin_clzll(before);
	owner_at[lane] = 0;
	__builtin_amdgcn_wave_barrier();
	const int e = excl - base;
	if (n > 0 && e > 0 && e < 64)
		owner_at[e] = lane + 1;
	__builtin_amdgcn_wave_barrier();
	const int mark = (lane == 0) ? o0 + 1 : owner_at[lane];
	const uint64_t marks = __ballot(mark != 0);
	const uint64_t upto_me = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
	const int at = 63 - (int)__builtin_clzll(marks & upto_me);
	return __shfl(mark, at) - 1;
}

// The same walk for the rays of a whole wave, with the candidate tests shared
// out over all 64 lanes.
//
// One lane walking its own cell's triangle list leaves most of the wave idle:
// lists differ in length, and the wave runs as long as the longest.  Here
// every round, each walking lane enters its next cell and counts its list
// (g.odc); the (ray, candidate) pairs of all lists are numbered by a prefix
// sum over lanes and screened 64 at a time, one per lane, with the owning
// ray's origin and direction fetched across lanes (cell_may_hit: material
// test and a conservative plane test; about 2 % of the pairs pass on a
// grating scan).  The pairs that pass are queued in LDS in pair order and
// tested in full 64 at a time.  The results that can hit (t > 0) are then
// taken in queue order, which is each owner's list order, and each owner
// applies the reference's running-minimum rule to its own, so hit, distance
// and tie choices are those of `propagate`.  (Testing every pair in full
// took about 20 % more of the step kernel.)
//
// Every lane of the wave must call it (lanes with walk == false take part in
// the shared tests only); a call from a partial wave falls back to
// `propagate` lane by lane.  `owner_at` is 192 ints of LDS for this wave:
// the owner table and the queue of pairs (triangle, owner lane).
//
// SHARED: every lane walks the same geometry.  Otherwise (lanes of different
// scans) the candidate tests read the owning lane's triangle tables.
//
// SAF (with SHARED): also return in `saf` the safety of each walking lane's
// start point (wall_safety of its first cell, and pair_safety of every
// triangle listed there, minimised per owner in the 64 ints of LDS after
// the owner table and queue), or -1 for a lane that does not walk.
template <bool SHARED, bool SAF = false>
__device__ inline int propagate_wave(const Geometry& g, bool walk, F3 start, F3 dir, float distance,
	int ignore_tri, int ignore_mat, float& out_dist, int* __restrict__ path, int stride, int* owner_at,
	float* saf = nullptr)
{
	out_dist = distance;
	if (SAF)
		*saf = -1.0f;
	if (__ballot(1) != ~0ull) {
		int r = -1;
		if (walk)
			r = propagate(g, start, dir, distance, ignore_tri, ignore_mat, out_dist, path, stride);
		return r;
	}
	const int lane = (int)__lane_id();
	F3 cur = start;
	float to_go = distance;
	uint64_t location = 1;
	uint32_t cells = 0;
	int result = -1;
	bool active = walk;
	bool first = true;   // the walk's first round: every walking lane in its start cell
	unsigned* s_saf = (unsigned*)owner_at + 192;
	while (__ballot(active) != 0ull) {
		int n = 0, off = 0, target = 0;
		float isect = 0.0f;
		if (active) {
			float ax, ay, az, aw;
			off = cell_find(g, cur, location, ax, ay, az, aw, path, stride, n);
			++cells;
			cell_exit(g, cur, dir, ax, ay, az, aw, isect, target);
			if (SAF && first)
				s_saf[lane] = __float_as_uint(fmaxf(wall_safety(g, cur, ax, ay, az, aw), 0.0f));
		}
		const int incl = wave_incl_sum(n);
		const int excl = incl - n;
		off -= excl;  // pair p of this lane tests od[off + p]
		const int total = __builtin_amdgcn_readlane(incl, 63);
		if constexpr (SHARED) {
			int* q_ti = owner_at + 64;
			int* q_own = owner_at + 128;
			int qn = 0;   // queued pairs (wave-uniform)
			// Second stage: test the queued pairs in full and fold the results
			// that can hit, in queue order, into their owners' running minimum.
			auto flush = [&]() {
				__builtin_amdgcn_wave_barrier();
				int ti = -1, own = 0;
				if (lane < qn) {
					ti = q_ti[lane];
					own = q_own[lane];
				}
				__builtin_amdgcn_wave_barrier();
				const F3 oc = { __shfl(cur.x, own), __shfl(cur.y, own), __shfl(cur.z, own) };
				const F3 odir = { __shfl(dir.x, own), __shfl(dir.y, own), __shfl(dir.z, own) };
				float t = -1.0f;
				if (lane < qn)
					t = intersect_ray(load_tri(g.tris, ti), oc, odir);
				uint64_t hits = __ballot(t > 0.0f);
				while (hits) {
					const int src = (int)__builtin_ctzll(hits);
					hits &= hits - 1ull;
					const float ts = __shfl(t, src);
					const int tis = __shfl(ti, src), os = __shfl(own, src);
					if (lane == os && ts <= isect + EPS) {
						isect = ts;
						target = tis;
					}
				}
			};
			for (int base = 0; base < total; base += 64) {
				const int p = base + lane;
				const int o = chunk_owner(base, excl, n, owner_at, lane);
				const F3 oc = { __shfl(cur.x, o), __shfl(cur.y, o), __shfl(cur.z, o) };
				const F3 odir = { __shfl(dir.x, o), __shfl(dir.y, o), __shfl(dir.z, o) };
				const int oign_tri = __shfl(ignore_tri, o), oign_mat = __shfl(ignore_mat, o);
				const float oto = __shfl(to_go, o);
				const int ooff = __shfl(off, o);
				bool pass = false;
				int ti = -1;
				if (p < total) {
					ti = g.od[ooff + p];
					const bool may = cell_may_hit(g, ooff + p, oc, odir, oign_mat, oto);
					pass = may && ti != oign_tri;
					if (SAF && first)
						atomicMin(s_saf + o, __float_as_uint(fmaxf(pair_safety(g, ooff + p, oc), 0.0f)));
				}
				const uint64_t pm = __ballot(pass);
				const int np = (int)__popcll(pm);
				if (qn + np > 64) {
					flush();
					qn = 0;
				}
				if (pass) {
					const int slot = qn + (int)__popcll(pm & ((1ull << lane) - 1ull));
					q_ti[slot] = ti;
					q_own[slot] = o;
				}
				qn += np;
			}
			if (qn > 0)
				flush();
			if (SAF && first) {
				__builtin_amdgcn_wave_barrier();
				if (walk)
					*saf = __uint_as_float(s_saf[lane]);
				__builtin_amdgcn_wave_barrier();
			}
		} else {
			// Lanes of different worlds (the multi-world kernel, 4 waves per
			// SIMD): every pair is tested in full; the screen's queue and
			// second load measured slower there.
			for (int base = 0; base < total; base += 64) {
				const int p = base + lane;
				const int o = chunk_owner(base, excl, n, owner_at, lane);
				const F3 oc = { __shfl(cur.x, o), __shfl(cur.y, o), __shfl(cur.z, o) };
				const F3 odir = { __shfl(dir.x, o), __shfl(dir.y, o), __shfl(dir.z, o) };
				const int oign_tri = __shfl(ignore_tri, o), oign_mat = __shfl(ignore_mat, o);
				const int ooff = __shfl(off, o);
				Geometry og = g;
				og.od = (const int*)shfl_ptr(g.od, o);
				og.tris = (const float4*)shfl_ptr(g.tris, o);
				float t = -1.0f;
				int ti = -1;
				if (p < total) {
					ti = og.od[ooff + p];
					if (ti != oign_tri)
						t = cell_candidate(og, ti, oc, odir, oign_mat);
				}
				// this lane's pairs within the chunk, lanes [lo, hi)
				uint64_t todo = 0;
				const uint64_t can_hit = __ballot(t > 0.0f);
				if (active) {
					const int lo = max(excl - base, 0), hi = min(incl - base, 64);
					if (lo < hi) {
						const uint64_t upto = (hi == 64) ? ~0ull : ((1ull << hi) - 1ull);
						todo = can_hit & upto & ~((1ull << lo) - 1ull);
					}
				}
				while (__ballot(todo != 0ull) != 0ull) {
					const int src = todo ? (int)__builtin_ctzll(todo) : lane;
					const float ts = __shfl(t, src);
					const int tis = __shfl(ti, src);
					if (todo) {
						if (ts <= isect + EPS) {
							isect = ts;
							target = tis;
						}
						todo &= todo - 1ull;
					}
				}
			}
		}
		first = false;
		if (active) {
			if (isect >= to_go) {
				out_dist = distance;
				result = -1;
				active = false;
			} else if (target >= 0) {
				out_dist = distance - to_go + isect;
				result = target;
				active = false;
			} else {
				to_go -= isect;
				cur = v_add(cur, v_mul(dir, isect));
				cell_next(location, target, dir);
				if (!(location > 1 && cells < MAX_CELLS)) {
					out_dist = distance;
					result = -1;
					active = false;
				}
			}
		}
	}
	return result;
}

// -------------------------------------------------------------- physics --
struct Particle { F3 pos, dir; float E; };

// physics/full_penn.h:50-57 (same body in elastic.h:54-62); the mean free
// path is multiplied by the mechanism's scale (exact when it is 1).  logE is
// m_log(E), computed once per step for both mechanisms.
__device__ __forceinline__ float mech_sample_path(const Table1& t, float scale, float logE, Rng& rng,
	float& tau)
{
	const float imfp = m_exp(t1_get(t, logE));
	tau = (1.0f / imfp) * scale;
	return -tau * m_log(rng_unit(rng));
}

// What the derivative estimator needs from a step's free-path race: the
// scaled mean free paths, the winning mechanism (0 none, 1 inelastic,
// 2 elastic) and the distance travelled; also log(E) of the step, which
// the elastic event uses again (E does not change in between).
struct PathInfo { float tau_el, tau_in, dist; int type; float logE; };

// physics/kieft/elastic.h:67-115; logE = m_log(p.E)
__device__ inline void elastic_execute(const Material& m, uint32_t fl, Particle& p, Rng& rng, float logE)
{
	Particle tp = p;
	float cos_theta, sin_theta;
	{
		const float x = logE;
		const float y = rng_unit(rng);
		cos_theta = clampr(t2_get(m.el_icdf, x, y), -1.0f, 1.0f);
		sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
	}
	tp.dir = normalised(tp.dir);
	const F3 normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
	tp.dir = v_add(v_mul(tp.dir, cos_theta), v_mul(normal_dir, sin_theta));
	if (tp.E < 200.0f) {
		if (fl & PH_EL_PHONON)
			tp.E = tp.E - fmin_std(50e-3f, m.phonon_loss);
	} else {
		if (fl & PH_EL_RECOIL)
			tp.E = tp.E - m.recoil_const * (1.0f - cos_theta) * tp.E;
	}
	p = tp;
}

// physics/full_penn.h:62-288
__device__ inline bool penn_execute(const Material& m, uint32_t fl, Particle& p, Rng& rng, Particle& se)
{
	Particle tp = p;
	float omega;
	{
		const float x = m_log(tp.E);
		const float y = rng_unit(rng);
		omega = m_exp(t2_get(m.in_icdf, x, y));
	}
	float binding;
	{
		const float x = m_log(tp.E);
		const float y = m_log(omega / tp.E);
		const float z = rng_unit(rng);
		binding = t3_get_rounddown(m.pn_ion, x, y, z);
	}
	tp.dir = normalised(tp.dir);

	if (binding <= 0.0f) {
		if (omega > m.band_gap) {
			float q;
			{
				const float x = m_log(tp.E);
				const float y = omega / tp.E;
				const float z = rng_unit(rng);
				q = t3_get(m.pn_q, x, y, z);
			}
			const F3 prim_normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
			F3 q_dir, q_normal_dir;
			{
				const float c = clampr((omega + q * q) / (2.0f * sqrtf(tp.E) * q), -1.0f, 1.0f);
				const float s = sqrtf(1.0f - c * c);
				q_dir = normalised(v_add(v_mul(tp.dir, c), v_mul(prim_normal_dir, s)));
				q_normal_dir = normalised(make_normal_vec(q_dir, rng_phi(rng)));
			}
			const float cpf = clampr((2.0f * tp.E - omega - q * q) / (2.0f * sqrtf(tp.E * (tp.E - omega))), -1.0f, 1.0f);
			const float spf = sqrtf(1.0f - cpf * cpf);
			tp.E = tp.E - omega;
			tp.dir = v_add(v_mul(tp.dir, cpf), v_mul(prim_normal_dir, spf));
			p = tp;
			if (!(fl & PH_PN_SE))
				return false;
			Particle sp;
			sp.pos = tp.pos;
			const float k_fermi = sqrtf(m.fermi);
			const float k_fermi_omega = sqrtf(m.fermi + omega);
			if (q < k_fermi_omega - k_fermi) {
				// plasmon decay (:141-194)
				const float f = m.fermi / omega;
				float x;
				if (f < 0.1f) {
					const float U = m_cbrt(rng_unit(rng));
					x = f * U * U;
				} else {
					const float lower = fmax_std(0.0f, f - 1.0f);
					const float upper = f;
					const float d1 = upper - lower;
					const float d2 = upper * upper - lower * lower;
					const float alpha = d2 / (d1 + d2);
					int rounds = 0;
					do {
						const float U1 = rng_unit(rng);
						const float U2 = rng_unit(rng);
						x = (U1 < alpha) ? sqrtf(d2 * U2 + lower * lower) : d1 * U2 + lower;
					} while (rng_unit(rng) > sqrtf(x * x + x) / (x + 0.5f) && ++rounds < MAX_REJECT);
				}
				sp.E = x * omega + omega;
				sp.dir = rng_uniform_vector(rng);
			} else if (q < k_fermi_omega + k_fermi) {
				// single excitation (:195-222)
				const float ki_parr = (omega - q * q) / (2.0f * q);
				const float kf_parr = ki_parr + q;
				const float ki_perp_min_sq = fmax_std(0.0f, m.fermi - kf_parr * kf_parr);
				const float ki_perp_max_sq = m.fermi - ki_parr * ki_parr;
				const float ki_perp_sq = fmax_std(0.0f,
					ki_perp_min_sq + (ki_perp_max_sq - ki_perp_min_sq) * rng_unit(rng));
				const F3 ki = v_add(s_mul(ki_parr, q_dir), s_mul(sqrtf(ki_perp_sq), q_normal_dir));
				sp.E = mag2(ki) + omega;
				sp.dir = v_add(ki, s_mul(q, q_dir));
			} else {
				return false;
			}
			se = sp;
			return true;
		}
		tp.E = tp.E - omega;   // sub-band-gap loss (:234-242)
		p = tp;
		return false;
	}

	// inner-shell excitation (:245-287)
	const F3 normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
	const float K = tp.E - m.fermi + 2.0f * binding;
	const float dK = binding + omega;
	const float cos_theta = sqrtf(dK / K);
	const float sin_theta = sqrtf(1.0f - saturater(cos_theta * cos_theta));
	F3 sdir = v_add(v_mul(tp.dir, cos_theta), v_mul(normal_dir, sin_theta));
	sdir = normalised(sdir);
	sdir = v_add(sdir, s_mul(sqrtf(binding / dK), rng_uniform_vector(rng)));
	sdir = normalised(sdir);
	bool made = false;
	if (fl & PH_PN_SE) {
		se.E = m.fermi + omega - binding;
		se.pos = tp.pos;
		se.dir = sdir;
		made = true;
	}
	tp.E = tp.E - omega;
	tp.dir = v_sub(tp.dir, s_mul(cos_theta, sdir));
	p = tp;
	return made;
}

// physics/kieft/inelastic.h:68-206
__device__ inline bool kieft_execute(const Material& m, uint32_t fl, Particle& p, Rng& rng, Particle& se)
{
	Particle tp = p;
	float omega0;
	{
		const float x = m_log(tp.E);
		const float y = rng_unit(rng);
		omega0 = m_exp(t2_get(m.in_icdf, x, y));
	}
	float binding;
	{
		const float x = m_log(omega0);
		const float y = rng_unit(rng);
		binding = t2_get_rounddown(m.ki_ion, x, y);
	}
	float omega;
	{
		const float omega_max = 0.5f * (tp.E + omega0 - m.fermi);
		float omega_min = omega0;
		float w0 = fmin_std(omega0 - 1.0f, fmax_std(0.0f, binding) - m.fermi);
		if (tp.E > 2.0f * omega0) {
			omega_min = 0.5f * (tp.E + omega0 - sqrtf(tp.E * (tp.E - 2.0f * omega0)));
			w0 = omega0;
		}
		const float U = rng_unit(rng);
		if ((w0 > 0.0f) && (omega_min > w0) && (omega_min < omega_max)) {
			const float f_min = 1.0f / w0 * m_log((omega_min - w0) / omega_min);
			const float f_max = 1.0f / w0 * m_log((omega_max - w0) / omega_max);
			omega = -w0 / m_expm1(w0 * (f_min * (1.0f - U) + f_max * U));
		} else {
			omega = omega0 * omega_max / (omega0 * (1.0f - U) + omega_max * U);
		}
	}
	if (binding < 0.0f) {
		if (m.band_gap < 0.0f) {
			// metals: no special case in the reference
		} else if (omega0 > m.band_gap) {
			binding = m.band_gap;
		} else {
			if (fl & PH_KI_OPTICAL) {
				tp.E = tp.E - omega0;
				p = tp;
			}
			return false;
		}
	}
	binding = fmax_std(0.0f, binding);
	tp.dir = normalised(tp.dir);
	const F3 normal_dir = normalised(make_normal_vec(tp.dir, rng_phi(rng)));
	const float K = tp.E - m.fermi + 2.0f * binding;
	const float dK = binding + omega;
	const float cos_theta = sqrtf(dK / K);
	const float sin_theta = sqrtf(1.0f - saturater(cos_theta * cos_theta));
	F3 sdir = v_add(v_mul(tp.dir, cos_theta), v_mul(normal_dir, sin_theta));
	if (fl & PH_KI_INST) {
		sdir = normalised(sdir);
		sdir = v_add(sdir, s_mul(sqrtf(binding / dK), rng_uniform_vector(rng)));
	}
	sdir = normalised(sdir);
	bool made = false;
	if (fl & PH_KI_SE) {
		se.E = m.fermi + omega - binding;
		se.pos = tp.pos;
		se.dir = sdir;
		made = true;
	}
	tp.E = tp.E - omega;
	if (fl & PH_KI_MOMCONS)
		tp.dir = v_sub(tp.dir, s_mul(cos_theta, sdir));
	p = tp;
	return made;
}

// physics/boundary_intersect.h:38-155
//
// With `g` given, it receives d log p / d dU of the random choices made
// (transmission, reflection, absorption; 0 if none), and mi / mo the
// materials with dU = b(mo) - b(mi).
__device__ inline int boundary_execute(const Material* mats, uint32_t fl, const Tri& tri,
	Particle& p, int& mat, Rng& rng, float* g = nullptr, int* g_mi = nullptr, int* g_mo = nullptr)
{
	if (g) *g = 0.0f;
	Particle tp = p;
	const F3 nd = normalised(tp.dir);
	const F3 n = normalised(cross(tri.e1, tri.e2));
	const float cos_theta = dot(n, nd);
	int mi, mo;
	if (cos_theta > 0.0f) { mi = tri.mat_in; mo = tri.mat_out; }
	else { mi = tri.mat_out; mo = tri.mat_in; }
	switch (mo) {
	case MAT_DETECTOR:
		return DETECTED;
	case MAT_DETECTOR_LT50:
		return (tp.E < 50.0f) ? DETECTED : ALIVE;
	case MAT_DETECTOR_GE50:
		return (tp.E >= 50.0f) ? DETECTED : ALIVE;
	case MAT_TERMINATOR:
		return TERMINATED;
	case MAT_MIRROR:
		tp.dir = v_sub(nd, v_mul(s_mul(2.0f, n), cos_theta));
		p = tp;
		return ALIVE;
	case MAT_NOP:
		return ALIVE;
	default:
		break;
	}
	float dU = 0.0f;
	if (mi >= 0) dU = dU - mats[mi].barrier;
	if (mo >= 0) dU = dU + mats[mo].barrier;
	if (g) { *g_mi = mi; *g_mo = mo; }
	if (tp.E * cos_theta * cos_theta + dU > 0.0f) {
		const float ec2 = tp.E * cos_theta * cos_theta;
		const float s = sqrtf(1.0f + dU / ec2);
		const float T = (fl & PH_BI_QUANTUM) ? 4.0f * s / ((1.0f + s) * (1.0f + s)) : 1.0f;
		if (rng_unit(rng) < T) {
			// d log T / d dU
			if (g && (fl & PH_BI_QUANTUM) && T < 1.0f)
				*g = (1.0f - s) / (2.0f * s * s * (1.0f + s) * ec2);
			if (fl & PH_BI_REFRACT)
				tp.dir = v_add(v_sub(nd, v_mul(n, cos_theta)), v_mul(s_mul(s, n), cos_theta));
			tp.E = tp.E + dU;
			mat = mo;
			p = tp;
			return ALIVE;
		}
		// d log (1 - T) / d dU
		if (g && (fl & PH_BI_QUANTUM) && T < 1.0f)
			*g = 2.0f / (s * dU);
	}
	if (fl & PH_BI_ABSORB) {
		if (dU < 0.0f) {
			const float P = m_exp(1.0f + 0.5f * tp.E / dU);
			const bool absorbed = rng_unit(rng) < P;
			if (g && P < 1.0f)
				*g = *g + (absorbed ? -0.5f * tp.E / (dU * dU) : P * 0.5f * tp.E / (dU * dU * (1.0f - P)));
			if (absorbed)
				return TERMINATED;
		}
	}
	tp.dir = v_sub(nd, v_mul(s_mul(2.0f, n), cos_theta));
	p = tp;
	return ALIVE;
}

// --------------------------------------------------------------- a step --
struct State {
	Particle p;
	int mat, last_tri;
	uint32_t tag, step, flags;
	uint64_t sid;
	// Safety left of the particle's last walk: no triangle of that walk's
	// start cell and no wall of that cell lies within this distance of the
	// particle (the step kernel's walk skip); < 0 when unknown.
	float safe;
};

struct World {
	Geometry g;
	const Material* mats;
	int nmats;
	uint32_t flags;
	float threshold;
	uint32_t key0, key1;
};

// Geometry and materials of one world of a multi-scan driver.
struct WorldDesc {
	Geometry g;
	const Material* mats;
	int nmats;
};

// init (drivers/gpu/gpu_driver.inl:403-480) in two halves around the walk.
//
// step_plan: domain and energy checks and path sampling.  Returns ALIVE or
// TERMINATED; for ALIVE, `distance` and `type` are the free path and its
// mechanism (0 in vacuum), `nd` the normalised direction.
__device__ __forceinline__ int step_plan(const World& W, const State& s, Rng& rng, F3& nd, float& distance,
	int& type, PathInfo& pinfo)
{
	pinfo.type = 0;
	pinfo.logE = 0.0f;
	type = 0;
	if (!in_domain(W.g, s.p.pos))
		return TERMINATED;
	distance = W.g.max_extent;
	float tau_el = 0.0f, tau_in = 0.0f;
	if (s.mat >= 0) {
		const Material& m = W.mats[s.mat];
		if (s.p.E < m.barrier + W.threshold)
			return TERMINATED;
		// core/scatter_list.h:101-107 with the reverse visiting order of
		// common/tuple.h:114-124: elastic first, strictly shorter wins.
		float dist = __builtin_inff();
		const float logE = m_log(s.p.E);
		pinfo.logE = logE;
		const float pe = mech_sample_path(m.el_imfp, m.el_scale, logE, rng, tau_el);
		if (pe < dist) { type = 2; dist = pe; }
		const float pi = mech_sample_path(m.in_imfp, m.in_scale, logE, rng, tau_in);
		if (pi < dist) { type = 1; dist = pi; }
		distance = dist;
	}
	nd = normalised(s.p.dir);
	pinfo.tau_el = tau_el;
	pinfo.tau_in = tau_in;
	return ALIVE;
}

// step_move: after the walk, move to the hit or to the end of the free path
// and name the event; stores the normalised direction.
__device__ __forceinline__ void step_move(State& s, F3 nd, float distance, int type, int hit, float hd,
	int& event, PathInfo& pinfo)
{
	s.p.dir = nd;
	if (hit < 0) {
		s.p.pos = v_add(s.p.pos, v_mul(nd, distance));
		event = type;
	} else {
		s.p.pos = v_add(s.p.pos, v_mul(nd, hd));
		s.last_tri = hit;
		event = EV_INTERSECT;
	}
	pinfo.dist = (hit < 0) ? distance : hd;
	pinfo.type = (s.mat >= 0) ? type : 0;
}

// The whole init for one lane: domain and energy checks, path sampling,
// propagate, move.  Returns ALIVE or TERMINATED and the event; `hit` is the
// intersected triangle.
__device__ inline int step_init(const World& W, State& s, Rng& rng, int& event, int& hit,
	int* path, int stride, PathInfo& pinfo)
{
	event = EV_NONE;
	hit = -1;
	F3 nd;
	float distance;
	int type;
	if (step_plan(W, s, rng, nd, distance, type, pinfo) != ALIVE)
		return TERMINATED;
	float hd;
	hit = propagate(W.g, s.p.pos, nd, distance, s.last_tri, s.mat, hd, path, stride);
	step_move(s, nd, distance, type, hit, hd, event, pinfo);
	return ALIVE;
}

// The event chosen by init (gpu_driver.inl:482-566).  For inelastic events
// the new parent state and secondary are computed into `s`/`se`; the caller
// commits them.
__device__ inline int step_event(const World& W, State& s, Rng& rng, int event, int hit,
	Particle& se, bool& se_made, float* g = nullptr, int* g_mi = nullptr, int* g_mo = nullptr)
{
	se_made = false;
	if (event == EV_INTERSECT) {
		const Tri tri = load_tri(W.g.tris, hit);
		return boundary_execute(W.mats, W.flags, tri, s.p, s.mat, rng, g, g_mi, g_mo);
	}
	if (event == EV_NONE)
		return ALIVE;
	s.last_tri = -1;
	const Material& m = W.mats[s.mat];
	if (event == EV_ELASTIC) {
		elastic_execute(m, W.flags, s.p, rng, m_log(s.p.E));
		return ALIVE;
	}
	se_made = (m.inel_model == 0) ? penn_execute(m, W.flags, s.p, rng, se)
		: kieft_execute(m, W.flags, s.p, rng, se);
	return ALIVE;
}

// The events of a step other than inelastic scattering: boundary crossing
// and elastic scattering (and nothing for EV_NONE).  The tracking kernel
// sends inelastic events to their own kernel, so its code holds neither the
// inelastic models nor their register pressure.  logE is m_log(s.p.E)
// from the step's path sampling.
__device__ inline int step_event_elastic_or_boundary(const World& W, State& s, Rng& rng, int event,
	int hit, float logE, float* g = nullptr, int* g_mi = nullptr, int* g_mo = nullptr)
{
	if (event == EV_INTERSECT) {
		const Tri tri = load_tri(W.g.tris, hit);
		return boundary_execute(W.mats, W.flags, tri, s.p, s.mat, rng, g, g_mi, g_mo);
	}
	if (event == EV_ELASTIC) {
		s.last_tri = -1;
		elastic_execute(W.mats[s.mat], W.flags, s.p, rng, logE);
	}
	return ALIVE;
}

}} // namespace nbl::dev
