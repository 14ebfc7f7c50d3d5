// HIP half of libnebula.so: device objects, the tracking kernels and the
// device side of the C ABI (include/nebula.h).
//
// Particle buffer design.  Particles in flight live in two structure-of-arrays
// buffers of `slots` entries used ping-pong.  In the buffer being read,
// survivors of the previous launch occupy [0, n_surv) and secondary electrons
// [slots - n_se, slots).  A launch of k_step moves every particle through up
// to K steps (init + event each), then appends it to the other buffer:
// survivors from the front (a wavefront ballot + prefix popcount + one
// atomic per wave), new secondaries from the back (made by k_inel, one
// atomic per block).  Dead particles
// are simply not copied, so the next launch reads a dense range.  A
// one-thread control kernel adopts the output counts and plans the refill of
// primaries from a device-resident queue; all sizes live in device memory, so
// an iteration never waits for the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../host/host.hpp"
#include "physics.hpp"

using namespace nbl;
using namespace nbl::dev;

#define HIP_OK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) \
	return fail(NBL_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); } while (0)

// ------------------------------------------------------------- buffers --
// One allocation per buffer of n entries: sid (8 bytes x n), then thirteen
// 4-byte fields x n (px py pz dx dy dz E mat last_tri tag step flags safe).  A
// buffer is one base pointer and its size, so a kernel holds four buffers
// in a few scalar registers instead of 13 pointers each.
struct Bufs {
	char* base;
	double* sc;             // scores, [parameter][entry], or null
	uint32_t sc_stride;     // entries of this buffer (n)
};
enum { F_PX, F_PY, F_PZ, F_DX, F_DY, F_DZ, F_E, F_MAT, F_LAST, F_TAG, F_STEP, F_FLAGS, F_SAFE, N_FIELDS };

__device__ __forceinline__ float* fld_f(const Bufs& b, int f)
{
	return (float*)(b.base + 8ull * b.sc_stride) + (size_t)f * b.sc_stride;
}
__device__ __forceinline__ uint32_t* fld_u(const Bufs& b, int f)
{
	return (uint32_t*)(b.base + 8ull * b.sc_stride) + (size_t)f * b.sc_stride;
}

// Likelihood-ratio scores carried by a particle (FP64 accumulation).
constexpr int MAX_PARAMS = 8;
constexpr int MAX_SCORED_MATERIALS = 16;
struct Score { double v[MAX_PARAMS]; };

// Parameter index of (material, kind) or -1; kinds 0 elastic MFP scale,
// 1 inelastic MFP scale, 2 barrier.
struct ScoreMap {
	int np;
	int8_t pidx[MAX_SCORED_MATERIALS][3];
};

struct Ctrl {
	uint32_t buf_in;        // buffer the next k_step reads
	uint32_t n_surv, n_se;  // particles in buf_in
	uint32_t se_limit;      // secondaries that fit the buffer's SE region
	uint32_t stepped;       // a k_step ran since the last k_ctrl
	uint32_t refill_n, refill_dst;
	uint32_t target;        // particles in flight to aim for
	uint32_t spill_n;       // secondaries parked on the spill stack
	uint32_t spill_take, spill_src;  // refill from the spill stack
	uint64_t refill_src;
	uint64_t queue_n, queue_next;
	// Counters that waves of the step kernels bump atomically.  Each sits on
	// a 128-byte line of its own (atomics to one line are serialised in L2),
	// and the two a wave always needs together share one 64-bit word, so
	// that one atomic returns both bases.
	alignas(128) uint64_t step_out;   // k_step: survivors (low 32) | inelastic queue (high 32)
	// k_inel, one atomic per block: secondaries made (low 32) | secondaries
	// dropped at birth, whose one step k_ctrl adds to `steps` (high 32)
	alignas(128) uint64_t se_ctr;
	alignas(128) uint32_t det_ctr;    // detected records written (may overshoot det_cap)
	alignas(128) uint32_t spill_ctr;  // spill stack top during a k_step (may overshoot)
	alignas(128) uint32_t dep_ctr;    // energy-deposit records written (may overshoot dep_cap)
	alignas(128) uint64_t steps;      // particle-steps (init calls on active particles)
};

// Counts held in the packed words.
__host__ __device__ inline uint32_t lo32(uint64_t v) { return (uint32_t)v; }
__host__ __device__ inline uint32_t hi32(uint64_t v) { return (uint32_t)(v >> 32); }
// survivors written by the last k_step + k_inel pair (k_inel keeps every
// queued particle), secondaries requested
__host__ __device__ inline uint32_t out_surv(const Ctrl& c) { return lo32(c.step_out) + hi32(c.step_out); }
__host__ __device__ inline uint32_t out_se(const Ctrl& c) { return lo32(c.se_ctr); }
// particle-steps so far, with those of the last launch's dropped secondaries
__host__ __device__ inline uint64_t total_steps(const Ctrl& c) { return c.steps + hi32(c.se_ctr); }

struct StepArgs {
	World W;
	const WorldDesc* worlds;    // per-world geometry and materials (MULTI kernels)
	Bufs buf[2];
	Bufs spill;
	Bufs inel;
	Ctrl* ctrl;
	nbl_detected* det;
	double* det_sc;         // scores of detected records, [record][parameter]
	nbl_deposit* dep;       // energy-deposit records, or null
	uint32_t dep_cap;
	ScoreMap map;
	uint32_t det_cap;
	uint32_t slots;
	uint32_t spill_cap;
	uint32_t K;
	// world-per-wave variant: the coming launch's particles grouped by
	// world (k_world_group); k_step's wave takes perm[i] for index i
	const uint32_t* perm;
	const uint32_t* perm_n;   // its padded length; perm_n[1]: one world, perm unused
};

__device__ __forceinline__ State load_state(const Bufs& b, uint32_t i)
{
	State s;
	s.p.pos = { fld_f(b, F_PX)[i], fld_f(b, F_PY)[i], fld_f(b, F_PZ)[i] };
	s.p.dir = { fld_f(b, F_DX)[i], fld_f(b, F_DY)[i], fld_f(b, F_DZ)[i] };
	s.p.E = fld_f(b, F_E)[i];
	s.mat = (int)fld_u(b, F_MAT)[i];
	s.last_tri = (int)fld_u(b, F_LAST)[i];
	s.tag = fld_u(b, F_TAG)[i];
	s.step = fld_u(b, F_STEP)[i];
	s.flags = fld_u(b, F_FLAGS)[i];
	s.safe = fld_f(b, F_SAFE)[i];
	s.sid = ((const uint64_t*)b.base)[i];
	return s;
}

__device__ __forceinline__ void store_state(const Bufs& b, uint32_t i, const State& s)
{
	fld_f(b, F_PX)[i] = s.p.pos.x; fld_f(b, F_PY)[i] = s.p.pos.y; fld_f(b, F_PZ)[i] = s.p.pos.z;
	fld_f(b, F_DX)[i] = s.p.dir.x; fld_f(b, F_DY)[i] = s.p.dir.y; fld_f(b, F_DZ)[i] = s.p.dir.z;
	fld_f(b, F_E)[i] = s.p.E;
	fld_u(b, F_MAT)[i] = (uint32_t)s.mat;
	fld_u(b, F_LAST)[i] = (uint32_t)s.last_tri;
	fld_u(b, F_TAG)[i] = s.tag;
	fld_u(b, F_STEP)[i] = s.step;
	fld_u(b, F_FLAGS)[i] = s.flags;
	fld_f(b, F_SAFE)[i] = s.safe;
	((uint64_t*)b.base)[i] = s.sid;
}

__device__ __forceinline__ void load_score(const Bufs& b, uint32_t i, int np, Score& sc)
{
#pragma unroll
	for (int q = 0; q < MAX_PARAMS; ++q)
		sc.v[q] = (q < np) ? b.sc[(size_t)q * b.sc_stride + i] : 0.0;
}

__device__ __forceinline__ void store_score(const Bufs& b, uint32_t i, int np, const Score& sc)
{
#pragma unroll
	for (int q = 0; q < MAX_PARAMS; ++q)
		if (q < np)
			b.sc[(size_t)q * b.sc_stride + i] = sc.v[q];
}

// score[q] += inc for a runtime q, without indexing registers dynamically
__device__ __forceinline__ void score_add(Score& sc, int q, float inc)
{
#pragma unroll
	for (int j = 0; j < MAX_PARAMS; ++j)
		if (j == q)
			sc.v[j] += (double)inc;
}

__device__ __forceinline__ int score_index(const ScoreMap& m, int mat, int kind)
{
	return (mat >= 0 && mat < MAX_SCORED_MATERIALS) ? (int)m.pidx[mat][kind] : -1;
}

// d log p / d theta of a step's free-path race in material `mat`: rates
// 1/tau, travelled distance d, winner `event` (or a boundary first).
__device__ __forceinline__ void score_free_path(Score& sc, const ScoreMap& map, const World& W,
	int mat, const PathInfo& pi, int event)
{
	if (pi.type == 0)
		return;
	const Material& m = W.mats[mat];
	const int pe = score_index(map, mat, 0), pn = score_index(map, mat, 1);
	if (pe >= 0)
		score_add(sc, pe, (pi.dist / pi.tau_el - ((event == EV_ELASTIC) ? 1.0f : 0.0f)) / m.el_scale);
	if (pn >= 0)
		score_add(sc, pn, (pi.dist / pi.tau_in - ((event == EV_INELASTIC) ? 1.0f : 0.0f)) / m.in_scale);
}

// d log p / d b for the barriers on either side of a boundary (dU = b_mo - b_mi)
__device__ __forceinline__ void score_barrier(Score& sc, const ScoreMap& map, int mi, int mo, float g)
{
	if (g == 0.0f)
		return;
	const int po = score_index(map, mo, 2), pi = score_index(map, mi, 2);
	if (po >= 0)
		score_add(sc, po, g);
	if (pi >= 0)
		score_add(sc, pi, -g);
}

// Slot allocation for the lanes of a wavefront with `want` set: one atomic
// per wave, ranks by prefix popcount of the ballot.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* ctr, bool want)
{
	const uint64_t m = __ballot(want);
	if (m == 0)
		return 0;
	const uint32_t lane = __lane_id();
	const int leader = __ffsll((unsigned long long)m) - 1;
	uint32_t base = 0;
	if ((int)lane == leader)
		base = atomicAdd(ctr, (uint32_t)__popcll(m));
	base = (uint32_t)__shfl((int)base, leader);
	return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Two slot kinds at once from a packed counter (kind a in the low 32 bits,
// kind b in the high 32 bits): one atomic per wave for both.
__device__ __forceinline__ void wave_alloc2(uint64_t* ctr, bool want_a, bool want_b, uint32_t& slot_a,
	uint32_t& slot_b)
{
	const uint64_t ma = __ballot(want_a), mb = __ballot(want_b);
	slot_a = slot_b = 0;
	if ((ma | mb) == 0)
		return;
	const uint32_t lane = __lane_id();
	const int leader = __ffsll((unsigned long long)(ma | mb)) - 1;
	uint64_t base = 0;
	if ((int)lane == leader)
		base = atomicAdd((unsigned long long*)ctr,
			((unsigned long long)__popcll(mb) << 32) | (unsigned long long)__popcll(ma));
	const uint32_t blo = (uint32_t)__shfl((int)lo32(base), leader);
	const uint32_t bhi = (uint32_t)__shfl((int)hi32(base), leader);
	const uint64_t below = (1ull << lane) - 1ull;
	slot_a = blo + (uint32_t)__popcll(ma & below);
	slot_b = bhi + (uint32_t)__popcll(mb & below);
}

// ------------------------------------------------------------- kernels --
// Adopt the previous launch's output and plan the refill from the queue.
// Secondaries that found no room in the particle buffer wait on the spill
// stack; they are taken back (newest first) before any new primary.
__global__ void k_ctrl(Ctrl* c, uint32_t slots, uint32_t spill_cap)
{
	if (threadIdx.x != 0 || blockIdx.x != 0)
		return;
	if (c->stepped) {
		c->steps += hi32(c->se_ctr);
		c->buf_in ^= 1u;
		c->n_surv = out_surv(*c);
		c->n_se = min(out_se(*c), c->se_limit);
		c->spill_n = min(c->spill_ctr, spill_cap);
		c->stepped = 0;
	}
	const uint32_t active = c->n_surv + c->n_se;
	uint32_t room = c->target > active ? c->target - active : 0u;
	room = min(room, slots - active);
	const uint32_t take = min(c->spill_n, room);
	room -= take;
	const uint64_t left = c->queue_n - c->queue_next;
	const uint32_t m = (uint32_t)min<uint64_t>(left, (uint64_t)room);
	c->spill_take = take;
	c->spill_src = c->spill_n - take;
	c->spill_n -= take;
	c->refill_src = c->queue_next;
	c->refill_dst = c->n_surv;
	c->refill_n = take + m;
	c->queue_next += m;
	c->n_surv += take + m;
	c->se_limit = slots - (c->n_surv + c->n_se);
	c->step_out = 0;
	c->se_ctr = 0;
	c->spill_ctr = c->spill_n;
	c->stepped = 1;
}

// Index i (0 .. n_in) of the coming step launch: survivors first, then the
// secondaries down from the top of the buffer.
__device__ __forceinline__ uint32_t in_slot(uint32_t i, uint32_t n_surv, uint32_t slots)
{
	return (i < n_surv) ? i : (slots - 1u - (i - n_surv));
}

// Grouping by world for the world-per-wave step kernel, in three kernels:
// k_world_group<false> counts the coming launch's particles per world;
// k_world_scan turns the counts into the worlds' starts, each a multiple
// of 64 so that every step wave holds one world, marks the padding after
// each world's range empty (~0u) and clears the counters;
// k_world_group<true> writes each index into its world's range, a wave's
// lanes of one world together, so that the step kernel's waves read runs
// of neighbouring slots.  wg holds counts [0, nw), starts [nw, 2 nw), fill
// levels [2 nw, 3 nw), the padded length of the index array at 3 nw, and
// at 3 nw + 1 a flag set when at most one world has particles: the
// launch's indices are then grouped as they stand, the placing pass
// returns at once and the step kernel reads index i as itself.
// Exclusive prefix of the counts into the starts (each padded to a
// multiple of 64, the padding marked ~0u), the padded length at wg[3 nw],
// the one-world flag at wg[3 nw + 1], counters and fill levels cleared for
// the next launch.
__device__ void world_scan(uint32_t nw, uint32_t* __restrict__ wg, uint32_t* __restrict__ perm)
{
	__shared__ uint32_t s_v[256];
	__shared__ uint32_t s_carry;
	if (threadIdx.x == 0)
		s_carry = 0;
	__syncthreads();
	uint32_t present = 0;   // worlds with particles
	for (uint32_t base = 0; base < nw; base += blockDim.x) {
		const uint32_t i = base + threadIdx.x;
		const uint32_t n = (i < nw) ? wg[i] : 0u;
		const uint32_t v = (n + 63u) & ~63u;
		present += (uint32_t)__syncthreads_count(n != 0u);
		s_v[threadIdx.x] = v;
		__syncthreads();
		for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
			const uint32_t add = (threadIdx.x >= o) ? s_v[threadIdx.x - o] : 0u;
			__syncthreads();
			s_v[threadIdx.x] += add;
			__syncthreads();
		}
		if (i < nw) {
			const uint32_t start = s_carry + s_v[threadIdx.x] - v;
			wg[nw + i] = start;
			wg[i] = 0;
			wg[2 * nw + i] = 0;
			for (uint32_t k = n; k < v; ++k)
				perm[start + k] = ~0u;
		}
		__syncthreads();
		if (threadIdx.x == blockDim.x - 1)
			s_carry += s_v[threadIdx.x];
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		wg[3 * nw] = s_carry;
		wg[3 * nw + 1] = (present <= 1u) ? 1u : 0u;
	}
}

// Each block takes a contiguous chunk of the indices (whole waves) and
// counts them per world in LDS first (one LDS atomic per world present in
// a wave), so that the counters in memory see one atomic per world present
// in a block: with one per wave, the waves' atomics on the few counters of
// the worlds in flight serialised in L2 (165 us per pass on the 64-scan
// batch).  PLACE: the block reserves its range of each world with those
// atomics and a second sweep writes the indices from LDS offsets.  Worlds
// beyond NW_LDS are counted with one atomic per wave instead.
constexpr uint32_t NW_LDS = 1024;
template <bool PLACE>
__global__ __launch_bounds__(256) void k_world_group(const Ctrl* c, Bufs b0, Bufs b1, uint32_t slots,
	uint32_t nw, uint32_t* __restrict__ wg, uint32_t* __restrict__ perm)
{
	__shared__ uint32_t s_h[NW_LDS];
	if (PLACE && wg[3 * nw + 1])
		return;   // one world: nothing to place
	const uint32_t n_surv = c->n_surv, n_in = n_surv + c->n_se;
	const Bufs& b = c->buf_in ? b1 : b0;
	const uint32_t lane = __lane_id();
	const bool lds = nw <= NW_LDS;
	// this block's chunk: whole waves
	const uint32_t per = ((n_in + gridDim.x - 1) / gridDim.x + 63u) & ~63u;
	const uint32_t lo = blockIdx.x * per, hi = min(n_in, lo + per);
	if (lds) {
		for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x)
			s_h[w] = 0;
		__syncthreads();
	}
	// sweep 0 counts; with PLACE, sweep 1 writes
	for (int sweep = 0; sweep < (PLACE ? 2 : 1); ++sweep) {
		if (sweep == 1) {
			__syncthreads();
			for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x)
				if (s_h[w])
					s_h[w] = wg[nw + w] + atomicAdd(&wg[2 * nw + w], s_h[w]);
			__syncthreads();
		}
		for (uint32_t w0 = lo + (threadIdx.x & ~63u); w0 < hi; w0 += blockDim.x) {
			const uint32_t i = w0 + lane;
			const bool valid = i < hi;
			uint32_t world = 0;
			if (valid)
				world = min(fld_u(b, F_FLAGS)[in_slot(i, n_surv, slots)] >> FL_WORLD_SHIFT, nw - 1u);
			uint64_t left = __ballot(valid);
			while (left) {
				const int first = __builtin_ctzll(left);
				const uint32_t w = __builtin_amdgcn_readlane(world, first);
				const uint64_t m = __ballot(valid && world == w);
				left &= ~m;
				const uint32_t cnt = (uint32_t)__popcll(m);
				if (sweep == 0) {
					if ((int)lane == first) {
						if (lds)
							atomicAdd(&s_h[w], cnt);
						else
							atomicAdd(&wg[w], cnt);
					}
				} else {
					uint32_t base = 0;
					if ((int)lane == first)
						base = atomicAdd(&s_h[w], cnt);
					base = __builtin_amdgcn_readlane(base, first);
					if ((m >> lane) & 1ull)
						perm[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
				}
			}
		}
	}
	if (!PLACE) {
		if (lds) {
			__syncthreads();
			for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x)
				if (s_h[w])
					atomicAdd(&wg[w], s_h[w]);
		}
	}
}

// One block: the counts into starts (a separate launch: a count pass
// whose last block did this needed a device-scope fence in every block,
// which on this multi-XCD GPU writes back the L2, 67 us against 9 + 5).
__global__ __launch_bounds__(256) void k_world_scan(uint32_t nw, uint32_t* __restrict__ wg,
	uint32_t* __restrict__ perm)
{
	world_scan(nw, wg, perm);
}

// The placing sweep for more than NW_LDS worlds: one atomic per world
// present in a wave.
__global__ __launch_bounds__(256) void k_world_place_wide(const Ctrl* c, Bufs b0, Bufs b1, uint32_t slots,
	uint32_t nw, uint32_t* __restrict__ wg, uint32_t* __restrict__ perm)
{
	if (wg[3 * nw + 1])
		return;   // one world: nothing to place
	const uint32_t n_surv = c->n_surv, n_in = n_surv + c->n_se;
	const Bufs& b = c->buf_in ? b1 : b0;
	const uint32_t stride = gridDim.x * blockDim.x;
	const uint32_t lane = __lane_id();
	const uint32_t wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * blockDim.x + (threadIdx.x & ~63u));
	for (uint32_t w0 = wave0; w0 < n_in; w0 += stride) {
		const uint32_t i = w0 + lane;
		const bool valid = i < n_in;
		uint32_t world = 0;
		if (valid)
			world = min(fld_u(b, F_FLAGS)[in_slot(i, n_surv, slots)] >> FL_WORLD_SHIFT, nw - 1u);
		uint64_t left = __ballot(valid);
		while (left) {
			const int first = __builtin_ctzll(left);
			const uint32_t w = __builtin_amdgcn_readlane(world, first);
			const uint64_t m = __ballot(valid && world == w);
			left &= ~m;
			uint32_t base = 0;
			if ((int)lane == first)
				base = wg[nw + w] + atomicAdd(&wg[2 * nw + w], (uint32_t)__popcll(m));
			base = __builtin_amdgcn_readlane(base, first);
			if ((m >> lane) & 1ull)
				perm[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
		}
	}
}

// Move planned primaries from the queue into the particle buffer
// (gpu_driver.inl:367-389 / gpu_particle_manager.inl:342-353: vacuum
// material, no last triangle).  Without a tag array a primary's tag is its
// queue index; without a stream-id array its stream id is sid_base + index.
__global__ __launch_bounds__(256) void k_refill(Ctrl* c, Bufs b0, Bufs b1, Bufs spill,
	const nbl_particle28* __restrict__ q, const uint32_t* __restrict__ qtag,
	const uint64_t* __restrict__ qsid, const uint16_t* __restrict__ qworld, uint32_t np,
	uint64_t sid_base)
{
	const uint32_t n = c->refill_n;
	const uint32_t stride = gridDim.x * blockDim.x;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
	const Bufs& b = c->buf_in ? b1 : b0;
	const uint32_t dst = c->refill_dst + i;
	const uint32_t take = c->spill_take;
	if (i < take) {
		store_state(b, dst, load_state(spill, c->spill_src + i));
		if (b.sc)
			for (uint32_t q = 0; q < np; ++q)
				b.sc[(size_t)q * b.sc_stride + dst] = spill.sc[(size_t)q * spill.sc_stride + c->spill_src + i];
		continue;
	}
	const uint64_t src = c->refill_src + (i - take);
	const nbl_particle28 p = q[src];
	State s;
	s.p.pos = { p.pos[0], p.pos[1], p.pos[2] };
	s.p.dir = { p.dir[0], p.dir[1], p.dir[2] };
	s.p.E = p.kin_energy;
	s.mat = MAT_VACUUM;
	s.last_tri = -1;
	s.tag = qtag ? qtag[src] : (uint32_t)src;
	s.step = 0;
	s.flags = qworld ? ((uint32_t)qworld[src] << FL_WORLD_SHIFT) : 0u;
	s.sid = qsid ? qsid[src] : sid_base + src;
	s.safe = -1.0f;
	store_state(b, dst, s);
	if (b.sc)
		for (uint32_t q = 0; q < np; ++q)
			b.sc[(size_t)q * b.sc_stride + dst] = 0.0;
	}
}

// Particle-step count: wave reduction, one atomic per wave.
__device__ __forceinline__ void count_steps(Ctrl* C, uint32_t my_steps)
{
	uint32_t v = my_steps;
	for (int off = 32; off > 0; off >>= 1)
		v += (uint32_t)__shfl_xor((int)v, off);
	if (__lane_id() == 0 && v)
		atomicAdd((unsigned long long*)&C->steps, (unsigned long long)v);
}

// Place a new secondary: the output buffer's SE region, else the spill
// stack; returns false when both are full.  With scores, the secondary
// starts from its parent's score.
template <bool SC>
__device__ __forceinline__ bool place_secondary(const StepArgs& a, Ctrl* C, const Bufs& out,
	uint32_t se_limit, bool made, uint32_t sslot, const Particle& se, const State& parent, const Score& psc)
{
	bool ok = true;
	if (made) {
		State ns;
		ns.p = se;
		ns.mat = parent.mat;
		ns.last_tri = -1;
		ns.tag = parent.tag;
		ns.step = 0;
		ns.flags = parent.flags & FL_WORLD_MASK;
		ns.sid = se_sid(parent.sid, parent.step);
		// it starts where its parent stands, so the parent's safety holds
		ns.safe = parent.safe;
		const bool spill = sslot >= se_limit;
		const uint32_t pslot = wave_alloc(&C->spill_ctr, spill);
		if (!spill) {
			store_state(out, a.slots - 1u - sslot, ns);
			if (SC) store_score(out, a.slots - 1u - sslot, a.map.np, psc);
		} else if (pslot < a.spill_cap) {
			store_state(a.spill, pslot, ns);
			if (SC) store_score(a.spill, pslot, a.map.np, psc);
		} else {
			ok = false;
		}
	}
	return ok;
}

// Up to K steps per particle: init (domain, threshold, path sampling, octree
// walk, move) and the boundary or elastic event.  A particle whose event is
// inelastic stops and joins the inelastic queue, so that the long inelastic
// code runs in k_inel on waves where every lane needs it (about 7% of the
// steps of a 300 eV scan would otherwise put it into nearly every wave).
//
// Grid-stride over the input particles by whole waves; each lane steps its
// particle up to K times, stopping early when it leaves (detected,
// terminated, inelastic event), and writes it out once: detected record,
// inelastic queue or survivor buffer, with one ballot per kind and one
// atomic per wave for the last two together.  The wave steps in lockstep because the octree walk is
// shared across it (propagate_wave): a lane whose particle has stopped still
// tests triangles for the others.  (Lanes that took new work after every step
// or every few steps were measured slower: the extra ballots and atomics cost
// more than the idle lanes they saved; so did lanes that kept a stopped
// particle until half the wave was idle and then refilled together from the
// wave's later particles, with a smaller grid to give them some.)
//
// SC: also accumulate each particle's likelihood-ratio score (free-path
// race and boundary choices) and write it with the particle.
// MULTI: particles belong to different worlds (flags bits 16..31); each lane
// takes its particle's geometry and materials from the world table.
// An energy-deposit record (energydeposit_cpu_driver.h:55-66): position
// after the event, energy before, energy lost; slots by ballot + atomic.
__device__ __forceinline__ void deposit(const StepArgs& a, Ctrl* C, bool rec, const State& s, float e0)
{
	const uint32_t slot = wave_alloc(&C->dep_ctr, rec);
	if (rec && slot < a.dep_cap) {
		nbl_deposit r;
		r.pos[0] = s.p.pos.x; r.pos[1] = s.p.pos.y; r.pos[2] = s.p.pos.z;
		r.energy = e0;
		r.deposit = e0 - s.p.E;
		r.tag = s.tag;
		r.sid = s.sid;
		a.dep[slot] = r;
	}
}

// WUNI: several worlds, one per wave.  A wave steps the particles of its
// first lane's world (its geometry and materials are wave-uniform, so the
// plain kernel's shared walk, walk skip and register budget apply); lanes
// holding another world's particle wait, unchanged, for a later launch.
// Trajectories depend only on each particle's own stream, so which launch
// steps a particle changes no result.
template <bool SC, bool MULTI, bool WUNI = false>
__device__ __forceinline__ void step_body(const StepArgs& a)
{
	static_assert(!(WUNI && (SC || MULTI)), "the world-per-wave variant is the plain one");
	// energy deposits are recorded by the general variants only
	constexpr bool DEP_OK = SC || MULTI;
	Ctrl* C = a.ctrl;
	const uint32_t bi = C->buf_in;
	const Bufs in = bi ? a.buf[1] : a.buf[0];
	const Bufs out = bi ? a.buf[0] : a.buf[1];
	const uint32_t n_surv = C->n_surv, n_in = n_surv + C->n_se;
	const uint32_t B = a.slots, det_cap = a.det_cap, K = a.K;
	const int np = a.map.np;
	World Wl = a.W;
	const World& W = (MULTI || WUNI) ? Wl : a.W;
	__shared__ int s_path[PATH_LEVELS * 256];
	__shared__ int s_owner[4 * 256];   // per wave: owner table, pair queue, safeties
	// In the plain variant a particle's stream id waits here during its steps
	// (each step reads it once to start its Philox stream): two fewer VGPRs
	// across the step loop, and the loop no longer spills (4 % faster).  The
	// multi-world variant has the registers and ran 3 % slower with it.
	constexpr bool SID_LDS = !SC && !MULTI;
	__shared__ uint64_t s_sid[SID_LDS ? 256 : 1];
	// With the stream id wait the particle's tag, step and flags (the walk
	// needs none of them): three VGPRs fewer across the step loop.
	__shared__ uint32_t s_aux[SID_LDS ? 3 * 256 : 1];
	uint32_t* const aux = s_aux + (SID_LDS ? threadIdx.x : 0u);

	int* path = s_path + threadIdx.x;
	int* owner_at = s_owner + 4 * (threadIdx.x & ~63u);
	// Steps whose free path cannot reach a triangle skip the walk: after a
	// walk that found no hit, `rem` is the start point's safety (no triangle
	// of its cell and no cell wall within that distance, propagate_wave)
	// less the distance moved since; a later free path no longer than `rem`
	// ends inside that ball, where the walk would find no hit and return the
	// full free path.  rem < 0: the next step walks.  It travels with the
	// particle (State::safe) from launch to launch, through the inelastic
	// queue (the event does not move it) and into its secondaries (they
	// start where it stands).  (Single-world variants; the multi-world walk
	// gives no safety and clears it.)
	constexpr bool SKIP = !MULTI;
	uint32_t my_steps = 0;
	const uint32_t stride = gridDim.x * blockDim.x;
	// Whole waves take consecutive particles and step in lockstep: the walk
	// shares its triangle tests across the wave (propagate_wave), so every
	// lane joins it each step, also lanes whose particle has stopped.
	const uint32_t wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * blockDim.x + (threadIdx.x & ~63u));
	// world-per-wave: index i names particle perm[i] (~0u: padding), or
	// particle i when the launch holds one world's particles only
	const bool ident = WUNI && a.perm_n[1] != 0u;
	const uint32_t n_loop = (WUNI && !ident) ? *a.perm_n : n_in;
	for (uint32_t w0 = wave0; w0 < n_loop; w0 += stride) {
		const uint32_t i = w0 + __lane_id();
		const uint32_t li = (WUNI && !ident) ? ((i < n_loop) ? a.perm[i] : ~0u) : i;
		const bool valid = (WUNI && !ident) ? li != ~0u : i < n_in;
		State s = {};
		Score sc;
		if (valid) {
			const uint32_t slot = (li < n_surv) ? li : (B - 1u - (li - n_surv));
			s = load_state(in, slot);
			if (SC) load_score(in, slot, np, sc);
			if (MULTI) {
				const WorldDesc& wd = a.worlds[s.flags >> FL_WORLD_SHIFT];
				Wl.g = wd.g;
				Wl.mats = wd.mats;
				Wl.nmats = wd.nmats;
			}
		}
		bool mine = valid;
		if (WUNI) {
			// lane 0 holds a particle (each world's range starts at a multiple
			// of 64 and has no padding before its last particle; with one
			// world, w0 < n_in): its world is the wave's, and no lane holds
			// another world's particle
			const uint32_t ww = __builtin_amdgcn_readfirstlane(s.flags >> FL_WORLD_SHIFT);
			// through the constant address space: scalar loads, so the
			// world's geometry and materials sit in SGPRs as the plain
			// kernel's arguments do (not in VGPRs)
			static_assert(sizeof(WorldDesc) % 4 == 0, "world descriptor in words");
			using CWord = const __attribute__((address_space(4))) uint32_t;
			CWord* src = (CWord*)(a.worlds + ww);
			WorldDesc wd;
			uint32_t* dst = reinterpret_cast<uint32_t*>(&wd);
#pragma unroll
			for (int k = 0; k < (int)(sizeof(WorldDesc) / 4); ++k)
				dst[k] = src[k];
			Wl.g = wd.g;
			Wl.mats = wd.mats;
			Wl.nmats = wd.nmats;
			mine = valid && (s.flags >> FL_WORLD_SHIFT) == ww;
		}
		if (SID_LDS) {
			s_sid[threadIdx.x] = s.sid;
			aux[0] = s.tag;
			aux[256] = s.step;
			aux[512] = s.flags;
		}
		bool run = mine, det = false, to_inel = false, dead = false;
		float rem = SKIP ? s.safe : -1.0f;
		for (uint32_t k = 0; k < K; ++k) {
			if (__ballot(run) == 0ull)
				break;
			int oc = ALIVE;
			bool walk = false;
			Rng rng;
			int mat0 = s.mat, type = 0;
			F3 nd = { 0.0f, 0.0f, 0.0f };
			float distance = 0.0f;
			PathInfo pinfo;
			if (run) {
				const uint32_t fl = SID_LDS ? aux[512] : s.flags;
				if (fl & FL_PEND_DET) {
					oc = DETECTED;
				} else if (fl & FL_PEND_INEL) {
					// an inelastic event that found no secondary slot earlier
					if (SID_LDS)
						aux[512] = fl & ~FL_PEND_INEL;
					else
						s.flags = fl & ~FL_PEND_INEL;
					to_inel = true;
				} else {
					rng_start(rng, W.key0, W.key1, SID_LDS ? s_sid[threadIdx.x] : s.sid, SID_LDS ? aux[256] : s.step);
					mat0 = s.mat;
					oc = step_plan(W, s, rng, nd, distance, type, pinfo);
					++my_steps;
					walk = (oc == ALIVE);
				}
			}
			const bool skip = SKIP && walk && distance <= rem;
			float hd, saf = -1.0f;
			int hit = propagate_wave<!MULTI, SKIP>(W.g, walk && !skip, s.p.pos, nd, distance, s.last_tri, s.mat,
				hd, path, 256, owner_at, &saf);
			if (SKIP && walk) {
				if (skip) {
					hit = -1;
					hd = distance;
				}
				rem = (hit >= 0) ? -1.0f : ((skip ? rem : saf) - distance) - W.g.step_eps;
			}
			if (walk) {
				int ev;
				step_move(s, nd, distance, type, hit, hd, ev, pinfo);
				if (SC)
					score_free_path(sc, a.map, W, mat0, pinfo, ev);
				if (ev == EV_INELASTIC) {
					to_inel = true;
				} else if (SC) {
					float g = 0.0f;
					int gmi = -1, gmo = -1;
					const float e0 = s.p.E;
					oc = step_event_elastic_or_boundary(W, s, rng, ev, hit, pinfo.logE, &g, &gmi, &gmo);
					if (ev == EV_INTERSECT)
						score_barrier(sc, a.map, gmi, gmo, g);
					if (DEP_OK && a.dep)
						deposit(a, C, ev == EV_ELASTIC && s.p.E != e0, s, e0);
				} else {
					const float e0 = s.p.E;
					oc = step_event_elastic_or_boundary(W, s, rng, ev, hit, pinfo.logE);
					if (DEP_OK && a.dep)
						deposit(a, C, ev == EV_ELASTIC && s.p.E != e0, s, e0);
				}
			}
			if (run) {
				if (oc == DETECTED) {
					det = true;
					run = false;
				} else if (oc == TERMINATED) {
					dead = true;
					run = false;
				} else if (to_inel) {
					run = false;
				} else if (SID_LDS) {
					aux[256] += 1u;
				} else {
					s.step++;
				}
			}
		}
		if (SID_LDS) {
			s.sid = s_sid[threadIdx.x];
			s.tag = aux[0];
			s.step = aux[256];
			s.flags = aux[512];
		}
		// the multi-world kernel moves particles without a safety: clear it
		s.safe = rem;
		bool keep = valid && !dead;
		const uint32_t dslot = wave_alloc(&C->det_ctr, det);
		if (det) {
			if (dslot < det_cap) {
				nbl_detected r;
				r.pos[0] = s.p.pos.x; r.pos[1] = s.p.pos.y; r.pos[2] = s.p.pos.z;
				r.dir[0] = s.p.dir.x; r.dir[1] = s.p.dir.y; r.dir[2] = s.p.dir.z;
				r.kin_energy = s.p.E;
				r.tag = s.tag;
				r.sid = s.sid;
				a.det[dslot] = r;
				if (SC) {
#pragma unroll
					for (int q = 0; q < MAX_PARAMS; ++q)
						if (q < np)
							a.det_sc[(size_t)dslot * np + q] = sc.v[q];
				}
				keep = false;
			} else {
				// no record slot: wait as a survivor and retry next launch
				s.flags |= FL_PEND_DET;
			}
		}
		const bool queue = keep && to_inel;
		const bool survive = keep && !to_inel;
		uint32_t oslot, qslot;
		wave_alloc2(&C->step_out, survive, queue, oslot, qslot);
		if (queue) {
			store_state(a.inel, qslot, s);
			if (SC) store_score(a.inel, qslot, np, sc);
		}
		if (survive) {
			store_state(out, oslot, s);
			if (SC) store_score(out, oslot, np, sc);
		}
	}
	count_steps(C, my_steps);
}

// The plain variant runs at 5 waves per SIMD (96 VGPRs, 6 of them spilled,
// 28 bytes of scratch per lane, hipcc -Rpass-analysis=kernel-resource-usage):
// measured faster than 6 waves (80 VGPRs, 2 % slower with the tag, step and
// flags in LDS, 11 % before) or 7.  The scored and multi-world variants
// would spill, so they are left free.
#ifndef NBL_PLAIN_WAVES
#define NBL_PLAIN_WAVES 5
#endif
template <bool SC, bool MULTI, bool WUNI = false>
__global__ __launch_bounds__(256) void k_step(StepArgs a)
{
	step_body<SC, MULTI, WUNI>(a);
}
template <>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NBL_PLAIN_WAVES, 8)))
void k_step<false, false, false>(StepArgs a)
{
	step_body<false, false, false>(a);
}
template <>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NBL_PLAIN_WAVES, 8)))
void k_step<false, false, true>(StepArgs a)
{
	step_body<false, false, true>(a);
}

// The inelastic events queued by k_step (drivers/gpu/gpu_driver.inl:505-540):
// forget the last triangle, run the inelastic model from draw 2 of the step's
// stream, place the secondary, and keep the particle for the next launch.
// The inelastic models carry no parameter of the score.
//
// Every queued particle comes back as a survivor, so queue entry i goes to
// survivor slot surv0 + i with no atomic.  Secondary slots are counted per
// block (ballot per wave, the four counts in LDS, one atomic per block):
// with one atomic per wave, the waves' atomics on one counter serialised in
// L2 and this kernel took 7 % more of the iteration.  The block loop runs
// whole blocks, so every thread reaches each barrier.
template <bool SC, bool MULTI>
__global__ __launch_bounds__(256) void k_inel(StepArgs a)
{
	Ctrl* C = a.ctrl;
	const uint32_t bi = C->buf_in;
	const Bufs out = bi ? a.buf[0] : a.buf[1];
	const uint32_t n = hi32(C->step_out), se_limit = C->se_limit, surv0 = lo32(C->step_out);
	const int np = a.map.np;
	World Wl = a.W;
	const World& W = MULTI ? Wl : a.W;
	const uint32_t stride = gridDim.x * blockDim.x;
	__shared__ uint32_t s_cnt[4], s_doa[4];
	__shared__ uint32_t s_base;
	const uint32_t wave = threadIdx.x >> 6;
	for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
		const uint32_t i = b0 + threadIdx.x;
		const bool valid = i < n;
		State s = {}, pre = {};
		Score sc;
		Particle se = {};
		bool made = false, doa = false;
		if (valid) {
			s = load_state(a.inel, i);
			if (MULTI) {
				const WorldDesc& wd = a.worlds[s.flags >> FL_WORLD_SHIFT];
				Wl.g = wd.g;
				Wl.mats = wd.mats;
				Wl.nmats = wd.nmats;
			}
			if (SC) load_score(a.inel, i, np, sc);
			pre = s;
			Rng rng;
			rng_start(rng, W.key0, W.key1, s.sid, s.step);
			rng_seek(rng, 2);
			step_event(W, s, rng, EV_INELASTIC, -1, se, made);
			// A secondary that its first step would end at once (outside the
			// domain, or below its material's barrier plus the threshold:
			// step_plan's checks, gpu_driver.inl:423-433) is not stored: it
			// leaves nothing behind, and its one step is counted (high half
			// of se_ctr; its parent's event always completes).
			if (made && (!in_domain(W.g, se.pos) || se.E < W.mats[s.mat].barrier + W.threshold)) {
				made = false;
				doa = true;
			}
		}
		const uint64_t m = __ballot(made);
		const uint64_t md = __ballot(doa);
		if (__lane_id() == 0) {
			s_cnt[wave] = (uint32_t)__popcll(m);
			s_doa[wave] = (uint32_t)__popcll(md);
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			const uint32_t tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
			const uint32_t totd = s_doa[0] + s_doa[1] + s_doa[2] + s_doa[3];
			s_base = (tot | totd) ? lo32(atomicAdd((unsigned long long*)&C->se_ctr,
				((unsigned long long)totd << 32) | tot)) : 0u;
		}
		__syncthreads();
		uint32_t sslot = s_base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
		for (uint32_t w = 0; w < wave; ++w)
			sslot += s_cnt[w];
		__syncthreads();   // s_cnt and s_base are rewritten by the next round
		if (valid) {
			bool done = true;
			if (place_secondary<SC>(a, C, out, se_limit, made, sslot, se, s, sc)) {
				s.step++;
			} else {
				// no room anywhere: retry the event from the same draws later
				s = pre;
				s.flags |= FL_PEND_INEL;
				done = false;
			}
			if ((SC || MULTI) && a.dep)
				deposit(a, C, done && s.p.E != pre.p.E, s, pre.p.E);
			store_state(out, surv0 + i, s);
			if (SC) store_score(out, surv0 + i, np, sc);
		}
	}
}

// Queue entries of one push: tag (tag_base + i, unless given), stream id
// sid_base + i and world number.
__global__ __launch_bounds__(256) void k_queue_ids(uint32_t* tag, uint64_t* sid, uint16_t* world, uint32_t n,
	int fill_tags, uint32_t tag_base, uint64_t sid_base, uint16_t w)
{
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		if (fill_tags)
			tag[i] = tag_base + i;
		sid[i] = sid_base + i;
		world[i] = w;
	}
}

__global__ void k_reset_det(Ctrl* c)
{
	if (threadIdx.x == 0 && blockIdx.x == 0)
		c->det_ctr = 0;
}

// Measurement aid: a plain grid-stride copy of 16-byte words.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		dst[i] = src[i];
}

// Test hook: batched propagate.
__global__ __launch_bounds__(256) void k_query(Geometry g, const float* pos, const float* dir,
	const float* dist, const int* ignore_tri, const int* ignore_mat, int* out_tri,
	float* out_dist, uint32_t n)
{
	__shared__ int s_path[PATH_LEVELS * 256];
	__shared__ int s_owner[4 * 256];   // per wave: owner table, pair queue, safeties
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	const bool valid = i < n;
	F3 p = { 0.0f, 0.0f, 0.0f }, d = { 1.0f, 0.0f, 0.0f };
	float dd = 0.0f;
	int it = -1, im = 0;
	if (valid) {
		p = { pos[3 * i], pos[3 * i + 1], pos[3 * i + 2] };
		d = { dir[3 * i], dir[3 * i + 1], dir[3 * i + 2] };
		dd = dist[i];
		it = ignore_tri[i];
		im = ignore_mat[i];
	}
	// the walk of the step kernel: whole waves, candidate tests shared
	float od;
	const int r = propagate_wave<true>(g, valid, p, d, dd, it, im, od, s_path + threadIdx.x, 256,
		s_owner + 4 * (threadIdx.x & ~63u));
	if (valid) {
		out_tri[i] = r;
		out_dist[i] = od;
	}
}

// Test hook: table lookups.
__global__ __launch_bounds__(256) void k_probe(Material m, int which, int mode, const float* x,
	const float* y, const float* z, float* out, uint32_t n)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float r = 0.0f;
	switch (which) {
	case 0: r = t1_get(m.el_imfp, x[i]); break;
	case 2: r = t1_get(m.in_imfp, x[i]); break;
	case 1: case 3: case 6: {
		const Table2& t = which == 1 ? m.el_icdf : which == 3 ? m.in_icdf : m.ki_ion;
		r = mode == 0 ? t2_get(t, x[i], y[i]) : mode == 1 ? t2_get_rounddown(t, x[i], y[i])
			: t2_get_nearest(t, x[i], y[i]);
		break;
	}
	case 4: case 5: {
		const Table3& t = which == 4 ? m.pn_q : m.pn_ion;
		r = mode == 0 ? t3_get(t, x[i], y[i], z[i]) : t3_get_rounddown(t, x[i], y[i], z[i]);
		break;
	}
	}
	out[i] = r;
}

// Test hook: one step of independent states.
__global__ __launch_bounds__(256) void k_step_states(World W, nbl_step_state* st, int* event,
	int* outcome, nbl_step_state* se_out, int* se_made_out, uint32_t n)
{
	__shared__ int s_path[PATH_LEVELS * 256];
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	nbl_step_state h = st[i];
	State s;
	s.p.pos = { h.pos[0], h.pos[1], h.pos[2] };
	s.p.dir = { h.dir[0], h.dir[1], h.dir[2] };
	s.p.E = h.kin_energy;
	s.mat = h.mat; s.last_tri = h.last_tri; s.tag = h.tag; s.step = h.step; s.flags = h.flags;
	s.sid = h.sid;
	s.safe = -1.0f;
	Rng rng;
	rng_start(rng, W.key0, W.key1, s.sid, s.step);
	int ev = EV_NONE, hit = -1;
	PathInfo pinfo;
	int oc = step_init(W, s, rng, ev, hit, s_path + threadIdx.x, 256, pinfo);
	Particle se;
	bool made = false;
	if (oc == ALIVE)
		oc = step_event(W, s, rng, ev, hit, se, made);
	event[i] = ev;
	outcome[i] = oc;
	se_made_out[i] = made ? 1 : 0;
	if (made) {
		nbl_step_state o;
		o.pos[0] = se.pos.x; o.pos[1] = se.pos.y; o.pos[2] = se.pos.z;
		o.dir[0] = se.dir.x; o.dir[1] = se.dir.y; o.dir[2] = se.dir.z;
		o.kin_energy = se.E;
		o.mat = s.mat; o.last_tri = -1; o.tag = s.tag; o.step = 0; o.flags = 0;
		o.sid = se_sid(s.sid, s.step);
		se_out[i] = o;
	}
	h.pos[0] = s.p.pos.x; h.pos[1] = s.p.pos.y; h.pos[2] = s.p.pos.z;
	h.dir[0] = s.p.dir.x; h.dir[1] = s.p.dir.y; h.dir[2] = s.p.dir.z;
	h.kin_energy = s.p.E;
	h.mat = s.mat; h.last_tri = s.last_tri; h.step = s.step + 1;
	st[i] = h;
}

// ------------------------------------------------------- host objects --
// ------------------------------------------ tracking without allocation --
// nbl_track_noalloc keeps its particle buffers, inelastic queue, spill stack,
// control block and material descriptors in one file-scope device array,
// `scratch`, instead of allocating them.  Each GPU holds its own copy (at
// its own address, looked up once per device when the device is opened)
// until the process ends; every call on that GPU reuses it, so calls that
// overlap on one GPU race.  Its size is fixed by the largest population it
// serves: NA_MAX_CAPACITY particles in flight (2 x that many buffer slots),
// NA_SPILL parked secondaries and NA_MAX_MATERIALS materials.
constexpr uint32_t NA_MAX_CAPACITY = 1u << 16;
constexpr uint32_t NA_SLOTS = 2u * NA_MAX_CAPACITY;
constexpr uint32_t NA_SPILL = 1u << 19;
constexpr int NA_MAX_MATERIALS = 64;
constexpr size_t na_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t na_bufs_bytes(size_t n) { return na_round(n * (8 + 4 * N_FIELDS)); }
constexpr size_t NA_OFF_MATS = na_round(sizeof(Ctrl));
constexpr size_t NA_OFF_BUF0 = NA_OFF_MATS + na_round(sizeof(Material) * NA_MAX_MATERIALS);
constexpr size_t NA_OFF_BUF1 = NA_OFF_BUF0 + na_bufs_bytes(NA_SLOTS);
constexpr size_t NA_OFF_INEL = NA_OFF_BUF1 + na_bufs_bytes(NA_SLOTS);
constexpr size_t NA_OFF_SPILL = NA_OFF_INEL + na_bufs_bytes(NA_SLOTS);
constexpr size_t NA_BYTES = NA_OFF_SPILL + na_bufs_bytes(NA_SPILL);
static_assert(NA_BYTES < (64u << 20), "the fixed tracking array stays below 64 MiB");

static __device__ __align__(256) char scratch[NA_BYTES];

// Start a run in the fixed array: a zeroed control block with the
// population target and the number of queued primaries (a kernel rather
// than a host copy, so nothing is staged in host memory).
__global__ void k_na_init(Ctrl* c, uint32_t target, uint64_t queue_n)
{
	if (threadIdx.x != 0 || blockIdx.x != 0)
		return;
	Ctrl z = {};
	z.target = target;
	z.queue_n = queue_n;
	*c = z;
}

struct nbl_device {
	int ordinal;
	int n_cu;           // compute units, for persistent grids
	hipStream_t stream;
	char* scratch;      // this device's address of `scratch`
};

struct nbl_geometry {
	nbl_device* dev;
	int* od;
	float4* tris;
	float4* scr;
	int2* odc;
	Geometry g;
	int max_material;   // largest physical material index named by a triangle
};

struct nbl_material {
	nbl_device* dev;
	float* tables[7];
	Material m;
};

struct nbl_driver {
	nbl_device* dev;
	hipStream_t stream;
	World W;
	Material* d_mats;
	uint32_t capacity, slots, K, det_cap;
	Bufs buf[2];
	void* buf_mem[2];
	Bufs spill;
	void* spill_mem;
	uint32_t spill_cap;
	Bufs inel;
	void* inel_mem;
	Ctrl* ctrl;
	Ctrl* h_ctrl;
	nbl_detected* det;
	nbl_particle28* q_part;
	uint32_t* q_tag;
	uint64_t* q_sid;
	uint16_t* q_world;
	uint64_t q_cap, q_n;
	std::vector<WorldDesc> h_worlds;   // world 0 = the creation geometry
	std::vector<Material*> world_mats; // device material arrays of worlds 1..
	std::vector<std::vector<Material>> h_world_mats;   // their host copies
	WorldDesc* d_worlds;
	uint32_t* wgrp;         // world grouping: counts, starts, fill levels (3 per world)
	uint32_t* perm;         // world grouping: the step launch's indices by world
	uint32_t next_tag;
	ScoreMap map;           // map.np == 0: no scores
	double* sc_mem[4];      // score arrays of buf[0], buf[1], spill, inel
	double* det_sc;
	std::vector<Material> h_mats;   // host copy of d_mats
	uint32_t blocks_per_cu = 64;    // k_step grid: blocks of 256 per compute unit
	uint32_t target = 0;            // particles in flight to aim for
	nbl_deposit* dep = nullptr;     // energy-deposit records (enabled when non-null)
	uint64_t dep_cap = 0;
	uint32_t dep_per_particle = 0;  // K + 1 when the buffer was sized
	bool use_graph = false;         // replay iterations from captured graphs
	uint64_t epoch = 0;             // bumped when a launch argument changes
	struct Graph { uint32_t n; uint64_t epoch; hipGraph_t g; hipGraphExec_t exec; };
	std::vector<Graph> graphs;
	uint32_t spill_growths = 0;     // doublings of the spill stack
	bool timing;
	std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used, ev_free;
	double step_ms;
	uint64_t step_launches;
	// pinned staging of detected records (and scores) for nbl_driver_run
	nbl_detected* h_stage = nullptr;
	double* h_stage_sc = nullptr;
	size_t stage_sc_n = 0;
	hipEvent_t stage_ev = nullptr;
};

namespace {

int set_device(const nbl_device* dev)
{
	HIP_OK(hipSetDevice(dev->ordinal));
	return NBL_OK;
}

template <typename T>
int dev_alloc(T** p, size_t count)
{
	*p = nullptr;
	if (count == 0)
		count = 1;
	hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
	if (e != hipSuccess) {
		*p = nullptr;
		return fail(NBL_ERR_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
	}
	return NBL_OK;
}

Table1 tab1(const nbl_table& t, const float* d) { return { d, t.dims[0], t.min[0], t.step[0] }; }
Table2 tab2(const nbl_table& t, const float* d)
{
	return { d, t.dims[0], t.dims[1], t.min[0], t.step[0], t.min[1], t.step[1] };
}
Table3 tab3(const nbl_table& t, const float* d)
{
	return { d, t.dims[0], t.dims[1], t.dims[2], t.min[0], t.step[0], t.min[1], t.step[1], t.min[2], t.step[2] };
}

size_t tcount(const nbl_table& t)
{
	if (!t.data || t.ndim <= 0)
		return 0;
	size_t c = 1;
	for (int a = 0; a < t.ndim; ++a)
		c *= (size_t)t.dims[a];
	return c;
}

uint32_t blocks_for(uint64_t n, uint32_t bs = 256)
{
	const uint64_t b = (n + bs - 1) / bs;
	return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

int alloc_bufs(Bufs& b, void*& mem, uint32_t slots)
{
	const size_t n = slots;
	const size_t bytes = n * (8 + 4 * N_FIELDS) + 256;
	char* p = nullptr;
	if (hipMalloc((void**)&p, bytes) != hipSuccess)
		return fail(NBL_ERR_ALLOC, "hipMalloc of the particle buffers failed");
	mem = p;
	b.base = p;
	b.sc = nullptr;
	b.sc_stride = slots;
	return NBL_OK;
}

int collect_timing(nbl_driver* d)
{
	for (auto& pr : d->ev_used) {
		float ms = 0;
		HIP_OK(hipEventElapsedTime(&ms, pr.first, pr.second));
		d->step_ms += ms;
		d->ev_free.push_back(pr);
	}
	d->ev_used.clear();
	return NBL_OK;
}

int read_ctrl(nbl_driver* d)
{
	HIP_OK(hipMemcpyAsync(d->h_ctrl, d->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, d->stream));
	HIP_OK(hipStreamSynchronize(d->stream));
	if (d->timing)
		return collect_timing(d);
	return NBL_OK;
}

uint32_t active_count(const Ctrl& c, uint32_t spill_cap)
{
	if (c.stepped)
		return out_surv(c) + std::min(out_se(c), c.se_limit) + std::min(c.spill_ctr, spill_cap);
	return c.n_surv + c.n_se + c.spill_n;
}

} // namespace

// ================================================================ C ABI ==
extern "C" {

int nbl_device_open(int ordinal, nbl_device** out)
{
	clear_error();
	if (!out)
		return fail(NBL_ERR_ARG, "null argument");
	int n = 0;
	HIP_OK(hipGetDeviceCount(&n));
	if (ordinal < 0 || ordinal >= n)
		return fail(NBL_ERR_ARG, "no such GPU ordinal " + std::to_string(ordinal));
	HIP_OK(hipSetDevice(ordinal));
	int n_cu = 0;
	HIP_OK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ordinal));
	// the address of the fixed tracking array on this device (host code
	// uses it, never the array's name, wherever it needs that memory)
	void* sp = nullptr;
	HIP_OK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
	if (!sp)
		return fail(NBL_ERR_DEVICE, "hipGetSymbolAddress returned no address for the tracking array");
	auto* d = new nbl_device;
	d->ordinal = ordinal;
	d->n_cu = n_cu > 0 ? n_cu : 1;
	d->scratch = (char*)sp;
	hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
	if (e != hipSuccess) {
		delete d;
		return fail(NBL_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
	}
	*out = d;
	return NBL_OK;
}

int nbl_device_close(nbl_device* d)
{
	clear_error();
	if (!d)
		return NBL_OK;
	hipSetDevice(d->ordinal);
	hipStreamSynchronize(d->stream);
	hipStreamDestroy(d->stream);
	delete d;
	return NBL_OK;
}

int nbl_device_synchronize(nbl_device* d)
{
	clear_error();
	if (!d)
		return fail(NBL_ERR_ARG, "null device");
	HIP_OK(hipSetDevice(d->ordinal));
	HIP_OK(hipStreamSynchronize(d->stream));
	return NBL_OK;
}

int nbl_geometry_upload(nbl_device* dev, const int32_t* od, int64_t n_od, const nbl_triangle44* tris,
	int64_t ntri, const float center[3], const float halfsize[3], nbl_geometry** out)
{
	clear_error();
	if (!dev || !od || n_od < 9 || !tris || ntri < 0 || !center || !halfsize || !out)
		return fail(NBL_ERR_ARG, "invalid geometry arguments");
	// Walk the whole table once: every child offset must point at a node
	// inside it and every leaf entry at a triangle, so that the traversal
	// kernel cannot read out of bounds.
	{
		// (node offset, depth); leaves may sit at depth 21 at most (the
		// builder's limit, which the 64-bit location code and the per-level
		// path cache of the walk rely on)
		std::vector<uint8_t> seen((size_t)n_od, 0);
		std::vector<std::pair<int64_t, int>> stack{ { 0, 0 } };
		seen[0] = 1;
		while (!stack.empty()) {
			const int64_t node = stack.back().first;
			const int depth = stack.back().second;
			stack.pop_back();
			if (node + 8 > n_od)
				return fail(NBL_ERR_FORMAT, "octree node out of range at " + std::to_string(node));
			if (depth + 1 > PATH_LEVELS)
				return fail(NBL_ERR_FORMAT, "octree deeper than 21 levels");
			for (int o = 0; o < 8; ++o) {
				const int64_t v = od[node + o];
				if (v > 0) {
					if (v >= n_od || seen[v])
						return fail(NBL_ERR_FORMAT, "bad octree child offset at " + std::to_string(node + o));
					seen[v] = 1;
					stack.push_back({ v, depth + 1 });
				} else {
					int64_t p = -v;
					if (p <= 0 || p >= n_od)
						return fail(NBL_ERR_FORMAT, "bad octree leaf offset at " + std::to_string(node + o));
					for (;; ++p) {
						if (p >= n_od)
							return fail(NBL_ERR_FORMAT, "unterminated octree leaf");
						if (od[p] < 0)
							break;
						if (od[p] >= ntri)
							return fail(NBL_ERR_FORMAT, "octree leaf names triangle " + std::to_string(od[p]));
					}
				}
			}
		}
	}
	int rc = set_device(dev);
	if (rc) return rc;
	auto* g = new nbl_geometry();
	g->dev = dev;
	for (int64_t i = 0; i < ntri; ++i)
		if (std::max(tris[i].material_in, tris[i].material_out) > 32767
			|| std::min(tris[i].material_in, tris[i].material_out) < -32768) {
			delete g;
			return fail(NBL_ERR_ARG, "triangle " + std::to_string(i) + ": material id outside the 16-bit range");
		}
	if ((rc = dev_alloc(&g->od, (size_t)n_od)) || (rc = dev_alloc(&g->tris, (size_t)ntri * 3))
		|| (rc = dev_alloc(&g->scr, (size_t)n_od * 2)) || (rc = dev_alloc(&g->odc, (size_t)n_od))) {
		hipFree(g->od);
		hipFree(g->tris);
		hipFree(g->scr);
		delete g;
		return rc;
	}
	// cnt[i]: number of triangle entries from od[i] up to the next negative
	// entry (the length of a leaf list when i is its start)
	std::vector<int32_t> cnt((size_t)n_od, 0);
	for (int64_t i = n_od - 1; i >= 0; --i)
		cnt[i] = (od[i] < 0 || i + 1 == n_od) ? 0 : cnt[i + 1] + 1;
	std::vector<float4> packed((size_t)ntri * 3), tscr((size_t)ntri * 2);
	for (int64_t i = 0; i < ntri; ++i) {
		const nbl_triangle44& t = tris[i];
		float ib, ob;
		std::memcpy(&ib, &t.material_in, 4);
		std::memcpy(&ob, &t.material_out, 4);
		packed[3 * i] = make_float4(t.r0[0], t.r0[1], t.r0[2], t.e1[0]);
		packed[3 * i + 1] = make_float4(t.e1[1], t.e1[2], t.e2[0], t.e2[1]);
		packed[3 * i + 2] = make_float4(t.e2[2], ib, ob, 0.0f);
		// screening record: e1 x e2 in the device's operation order
		// (cross() in physics.hpp; no contraction on either side)
		const float nx = t.e1[1] * t.e2[2] - t.e1[2] * t.e2[1];
		const float ny = t.e1[2] * t.e2[0] - t.e1[0] * t.e2[2];
		const float nz = t.e1[0] * t.e2[1] - t.e1[1] * t.e2[0];
		const float l1 = std::fabs(t.e1[0]) + std::fabs(t.e1[1]) + std::fabs(t.e1[2]);
		const float l2 = std::fabs(t.e2[0]) + std::fabs(t.e2[1]) + std::fabs(t.e2[2]);
		const uint32_t mats = ((uint32_t)t.material_in & 0xffffu) | ((uint32_t)t.material_out << 16);
		float mf;
		std::memcpy(&mf, &mats, 4);
		tscr[2 * i] = make_float4(t.r0[0], t.r0[1], t.r0[2], l1 * l2);
		tscr[2 * i + 1] = make_float4(nx, ny, nz, mf);
	}
	// the screening records in table order: every entry that names a
	// triangle (list entries; node offsets that happen to be below ntri get a
	// record nobody reads)
	std::vector<float4> scr((size_t)n_od * 2, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
	for (int64_t i = 0; i < n_od; ++i)
		if (od[i] >= 0 && od[i] < ntri) {
			scr[2 * i] = tscr[2 * (size_t)od[i]];
			scr[2 * i + 1] = tscr[2 * (size_t)od[i] + 1];
		}
	HIP_OK(hipMemcpy(g->od, od, sizeof(int32_t) * n_od, hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(g->tris, packed.data(), sizeof(float4) * packed.size(), hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(g->scr, scr.data(), sizeof(float4) * scr.size(), hipMemcpyHostToDevice));
	// child entries with their list lengths (leaf offsets are checked above
	// to lie in (0, n_od))
	std::vector<int2> odc((size_t)n_od);
	for (int64_t i = 0; i < n_od; ++i) {
		const int32_t v = od[i];
		odc[i] = make_int2(v, (v < 0 && (int64_t)-(int64_t)v < n_od) ? cnt[(size_t)-(int64_t)v] : 0);
	}
	HIP_OK(hipMemcpy(g->odc, odc.data(), sizeof(int2) * n_od, hipMemcpyHostToDevice));
	g->g.od = g->od;
	g->g.tris = g->tris;
	g->g.scr = g->scr;
	g->g.odc = g->odc;
	g->g.center = { center[0], center[1], center[2] };
	g->g.halfsize = { halfsize[0], halfsize[1], halfsize[2] };
	// octree.inl:298-300: 2 |halfsize|, CPU expression order
	const float h2 = halfsize[0] * halfsize[0] + halfsize[1] * halfsize[1] + halfsize[2] * halfsize[2];
	g->g.max_extent = 2 * std::sqrt(h2);
	g->g.ntri = (int)ntri;
	g->g.step_eps = 3e-7f * (std::fabs(center[0]) + halfsize[0] + std::fabs(center[1]) + halfsize[1] +
		std::fabs(center[2]) + halfsize[2]) + 1e-6f;
	g->max_material = -1;
	for (int64_t i = 0; i < ntri; ++i)
		g->max_material = std::max(g->max_material, std::max(tris[i].material_in, tris[i].material_out));
	*out = g;
	return NBL_OK;
}

int nbl_geometry_destroy(nbl_geometry* g)
{
	clear_error();
	if (!g)
		return NBL_OK;
	hipSetDevice(g->dev->ordinal);
	hipFree(g->od);
	hipFree(g->tris);
	hipFree(g->scr);
	hipFree(g->odc);
	delete g;
	return NBL_OK;
}

int nbl_material_upload(nbl_device* dev, const nbl_material_host* m, nbl_material** out)
{
	clear_error();
	if (!dev || !m || !out)
		return fail(NBL_ERR_ARG, "null argument");
	if (m->inel_model != NBL_INEL_PENN && m->inel_model != NBL_INEL_KIEFT)
		return fail(NBL_ERR_ARG, "unknown inelastic model " + std::to_string(m->inel_model));
	// Every table the kernels read must be present with the dimensions its
	// lookups assume (at least two entries per axis: the interpolation reads
	// index n - 2 and n - 1), so that no lookup reads outside it.
	const nbl_table* ts[7] = { &m->el_imfp, &m->el_icdf, &m->in_imfp, &m->in_icdf, &m->pn_q, &m->pn_ion, &m->ki_ion };
	static const char* const names[7] = { "elastic imfp", "elastic cos(theta) icdf", "inelastic imfp",
		"inelastic icdf", "Penn q icdf", "Penn binding icdf", "Kieft binding icdf" };
	const bool penn = m->inel_model == NBL_INEL_PENN;
	const int want[7] = { 1, 2, 1, 2, penn ? 3 : 0, penn ? 3 : 0, penn ? 0 : 2 };
	for (int k = 0; k < 7; ++k) {
		if (!want[k])
			continue;
		const nbl_table& t = *ts[k];
		if (!t.data || t.ndim != want[k])
			return fail(NBL_ERR_FORMAT, std::string(names[k]) + " table missing or not " +
				std::to_string(want[k]) + "-D");
		for (int a = 0; a < t.ndim; ++a)
			if (t.dims[a] < 2)
				return fail(NBL_ERR_FORMAT, std::string(names[k]) + " table needs at least 2 entries per axis");
	}
	int rc = set_device(dev);
	if (rc) return rc;
	auto* mm = new nbl_material();
	mm->dev = dev;
	auto release = [&](int code) {
		for (float*& t : mm->tables) { hipFree(t); t = nullptr; }
		delete mm;
		return code;
	};
	for (int k = 0; k < 7; ++k) {
		const size_t c = want[k] ? tcount(*ts[k]) : 0;
		mm->tables[k] = nullptr;
		if (!c)
			continue;
		if ((rc = dev_alloc(&mm->tables[k], c)))
			return release(rc);
		const hipError_t e = hipMemcpy(mm->tables[k], ts[k]->data, sizeof(float) * c, hipMemcpyHostToDevice);
		if (e != hipSuccess)
			return release(fail(NBL_ERR_DEVICE, std::string("material table upload: ") + hipGetErrorString(e)));
	}
	Material& d = mm->m;
	d.barrier = m->barrier;
	d.phonon_loss = m->phonon_loss;
	d.recoil_const = m->recoil_const;
	d.fermi = m->fermi;
	d.band_gap = m->band_gap;
	d.el_scale = 1.0f;
	d.in_scale = 1.0f;
	d.inel_model = m->inel_model;
	d.el_imfp = tab1(m->el_imfp, mm->tables[0]);
	d.el_icdf = tab2(m->el_icdf, mm->tables[1]);
	d.in_imfp = tab1(m->in_imfp, mm->tables[2]);
	d.in_icdf = tab2(m->in_icdf, mm->tables[3]);
	d.pn_q = tab3(m->pn_q, mm->tables[4]);
	d.pn_ion = tab3(m->pn_ion, mm->tables[5]);
	d.ki_ion = tab2(m->ki_ion, mm->tables[6]);
	*out = mm;
	return NBL_OK;
}

int nbl_material_destroy(nbl_material* m)
{
	clear_error();
	if (!m)
		return NBL_OK;
	hipSetDevice(m->dev->ordinal);
	for (float* t : m->tables)
		hipFree(t);
	delete m;
	return NBL_OK;
}

int nbl_driver_destroy(nbl_driver* d)
{
	clear_error();
	if (!d)
		return NBL_OK;
	hipSetDevice(d->dev->ordinal);
	hipStreamSynchronize(d->stream);
	for (auto& pr : d->ev_used) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
	for (auto& pr : d->ev_free) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
	hipFree(d->buf_mem[0]);
	hipFree(d->buf_mem[1]);
	hipFree(d->spill_mem);
	hipFree(d->inel_mem);
	hipFree(d->ctrl);
	hipHostFree(d->h_ctrl);
	hipHostFree(d->h_stage);
	hipHostFree(d->h_stage_sc);
	if (d->stage_ev) hipEventDestroy(d->stage_ev);
	hipFree(d->det);
	hipFree(d->d_mats);
	hipFree(d->q_part);
	hipFree(d->q_tag);
	hipFree(d->q_sid);
	hipFree(d->q_world);
	for (auto& gr : d->graphs) { hipGraphExecDestroy(gr.exec); hipGraphDestroy(gr.g); }
	for (Material* p : d->world_mats)
		hipFree(p);
	hipFree(d->d_worlds);
	hipFree(d->wgrp);
	hipFree(d->perm);
	hipFree(d->dep);
	for (double* p : d->sc_mem)
		hipFree(p);
	hipFree(d->det_sc);
	hipStreamDestroy(d->stream);
	delete d;
	return NBL_OK;
}

int nbl_driver_create(nbl_device* dev, uint32_t capacity, const nbl_geometry* geom,
	const nbl_material* const* mats, int32_t n_mats, uint32_t physics_flags,
	float energy_threshold, uint64_t seed, uint32_t steps_per_launch, nbl_driver** out)
{
	clear_error();
	if (!dev || !geom || !out || (n_mats > 0 && !mats) || n_mats < 0)
		return fail(NBL_ERR_ARG, "invalid driver arguments");
	if (capacity < 64 || capacity > (1u << 30))
		return fail(NBL_ERR_CAPACITY, "capacity must be in [64, 2^30]");
	if (steps_per_launch < 1 || steps_per_launch > 1024)
		return fail(NBL_ERR_ARG, "steps_per_launch must be in [1, 1024]");
	for (int i = 0; i < n_mats; ++i)
		if (!mats[i] || mats[i]->dev != dev)
			return fail(NBL_ERR_ARG, "material " + std::to_string(i) + " missing or on another device");
	// gpu_main.cu:313-335: every material the geometry names must be given
	if (geom->max_material >= n_mats)
		return fail(NBL_ERR_ARG, "not enough materials provided for this geometry: expected " +
			std::to_string(geom->max_material + 1) + ", got " + std::to_string(n_mats));
	int rc = set_device(dev);
	if (rc) return rc;
	auto* d = new nbl_driver();
	d->dev = dev;
	d->capacity = capacity;
	d->slots = 2 * capacity;
	d->K = steps_per_launch;
	d->det_cap = capacity;
	auto cleanup = [&](int code) { nbl_driver_destroy(d); return code; };
	hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
	if (e != hipSuccess) {
		delete d;
		return fail(NBL_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
	}
	if ((rc = alloc_bufs(d->buf[0], d->buf_mem[0], d->slots))) return cleanup(rc);
	if ((rc = alloc_bufs(d->buf[1], d->buf_mem[1], d->slots))) return cleanup(rc);
	// The spill stack starts at eight buffers' worth (60 bytes a slot: 2 GB
	// at 2^22 slots) and doubles when needed.  A scan of 300 eV primaries
	// filling 2^22 slots spills more than twice the buffer in its first
	// iterations; each doubling costs a copy and a reallocation with the
	// stream idle.
	d->spill_cap = std::max<uint32_t>(1024u, std::min<uint64_t>(8ull * d->slots, 1ull << 28));
	if ((rc = alloc_bufs(d->spill, d->spill_mem, d->spill_cap))) return cleanup(rc);
	if ((rc = alloc_bufs(d->inel, d->inel_mem, d->slots))) return cleanup(rc);
	if ((rc = dev_alloc(&d->ctrl, 1))) return cleanup(rc);
	if (hipHostMalloc((void**)&d->h_ctrl, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess)
		return cleanup(fail(NBL_ERR_ALLOC, "hipHostMalloc failed"));
	if ((rc = dev_alloc(&d->det, d->det_cap))) return cleanup(rc);
	// Pinned staging for the frame loop's record copies, allocated here as
	// the reference's driver allocates its pinned buffers
	// (drivers/gpu/gpu_driver.inl:81-83), with one device-to-host copy into
	// it on the driver's stream: the first large device-to-host copy on a
	// stream waits several ms inside hipMemcpyAsync before it starts, which
	// would otherwise idle the GPU at the end of the first frame.
	if (hipHostMalloc((void**)&d->h_stage, sizeof(nbl_detected) * d->det_cap, hipHostMallocDefault) != hipSuccess) {
		d->h_stage = nullptr;
		return cleanup(fail(NBL_ERR_ALLOC, "pinned staging for detected records"));
	}
	if (hipMemcpyAsync(d->h_stage, d->det, sizeof(nbl_detected) * std::min<size_t>(d->det_cap, 1u << 16),
			hipMemcpyDeviceToHost, d->stream) != hipSuccess || hipStreamSynchronize(d->stream) != hipSuccess)
		return cleanup(fail(NBL_ERR_DEVICE, "device-to-host copy into the staging buffer failed"));
	if ((rc = dev_alloc(&d->d_mats, (size_t)std::max(1, n_mats)))) return cleanup(rc);
	std::vector<Material> hm((size_t)std::max(1, n_mats));
	for (int i = 0; i < n_mats; ++i)
		hm[i] = mats[i]->m;
	if (hipMemcpy(d->d_mats, hm.data(), sizeof(Material) * hm.size(), hipMemcpyHostToDevice) != hipSuccess)
		return cleanup(fail(NBL_ERR_DEVICE, "material descriptor upload failed"));
	d->h_mats = hm;
	d->map.np = 0;
	std::memset(d->map.pidx, -1, sizeof(d->map.pidx));
	Ctrl c;
	std::memset(&c, 0, sizeof(c));
	c.target = capacity;
	d->target = capacity;
	if (hipMemcpy(d->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice) != hipSuccess)
		return cleanup(fail(NBL_ERR_DEVICE, "control block upload failed"));
	d->W.g = geom->g;
	d->W.mats = d->d_mats;
	d->W.nmats = n_mats;
	d->W.flags = physics_flags;
	d->W.threshold = energy_threshold;
	d->W.key0 = (uint32_t)seed;
	d->W.key1 = (uint32_t)(seed >> 32);
	WorldDesc w0;
	w0.g = geom->g;
	w0.mats = d->d_mats;
	w0.nmats = n_mats;
	d->h_worlds.push_back(w0);
	if ((rc = dev_alloc(&d->d_worlds, 1))) return cleanup(rc);
	if (hipMemcpy(d->d_worlds, &w0, sizeof(WorldDesc), hipMemcpyHostToDevice) != hipSuccess)
		return cleanup(fail(NBL_ERR_DEVICE, "world table upload failed"));
	*out = d;
	return NBL_OK;
}

int nbl_driver_add_world(nbl_driver* d, const nbl_geometry* geom, const nbl_material* const* mats,
	int32_t n_mats, int32_t* world)
{
	clear_error();
	if (!d || !geom || !world || (n_mats > 0 && !mats) || n_mats < 0)
		return fail(NBL_ERR_ARG, "invalid world arguments");
	if (d->h_worlds.size() >= 65536)
		return fail(NBL_ERR_CAPACITY, "at most 65536 worlds per driver");
	if (geom->dev != d->dev)
		return fail(NBL_ERR_ARG, "geometry on another device");
	for (int i = 0; i < n_mats; ++i)
		if (!mats[i] || mats[i]->dev != d->dev)
			return fail(NBL_ERR_ARG, "material " + std::to_string(i) + " missing or on another device");
	if (geom->max_material >= n_mats)
		return fail(NBL_ERR_ARG, "not enough materials provided for this geometry: expected " +
			std::to_string(geom->max_material + 1) + ", got " + std::to_string(n_mats));
	int rc = set_device(d->dev);
	if (rc) return rc;
	Material* dm = nullptr;
	if ((rc = dev_alloc(&dm, (size_t)std::max(1, n_mats))))
		return rc;
	std::vector<Material> hm((size_t)std::max(1, n_mats));
	for (int i = 0; i < n_mats; ++i)
		hm[i] = mats[i]->m;
	if (hipMemcpy(dm, hm.data(), sizeof(Material) * hm.size(), hipMemcpyHostToDevice) != hipSuccess) {
		hipFree(dm);
		return fail(NBL_ERR_DEVICE, "material descriptor upload failed");
	}
	d->world_mats.push_back(dm);
	d->h_world_mats.push_back(hm);
	WorldDesc w;
	w.g = geom->g;
	w.mats = dm;
	w.nmats = n_mats;
	d->h_worlds.push_back(w);
	// the table is read by kernels in flight: replace it only when idle
	HIP_OK(hipStreamSynchronize(d->stream));
	WorldDesc* nw = nullptr;
	if ((rc = dev_alloc(&nw, d->h_worlds.size())))
		return rc;
	HIP_OK(hipMemcpy(nw, d->h_worlds.data(), sizeof(WorldDesc) * d->h_worlds.size(), hipMemcpyHostToDevice));
	hipFree(d->d_worlds);
	d->d_worlds = nw;
	// the world grouping's counters (cleared here, then by k_world_scan)
	// and its index array
	uint32_t* ng = nullptr;
	if ((rc = dev_alloc(&ng, 3 * d->h_worlds.size() + 2)))
		return rc;
	HIP_OK(hipMemset(ng, 0, sizeof(uint32_t) * (3 * d->h_worlds.size() + 2)));
	hipFree(d->wgrp);
	d->wgrp = ng;
	// every world's range padded to a multiple of 64
	uint32_t* np = nullptr;
	if ((rc = dev_alloc(&np, (size_t)d->slots + 64 * d->h_worlds.size())))
		return rc;
	hipFree(d->perm);
	d->perm = np;
	d->epoch++;
	*world = (int32_t)d->h_worlds.size() - 1;
	return NBL_OK;
}

int nbl_driver_push_world(nbl_driver* d, const nbl_particle28* particles, const uint32_t* tags,
	uint32_t n, uint64_t sid_base, int32_t world, uint32_t* pushed)
{
	clear_error();
	if (!d || (n && !particles))
		return fail(NBL_ERR_ARG, "invalid push arguments");
	if (world < 0 || world >= (int32_t)d->h_worlds.size())
		return fail(NBL_ERR_ARG, "no such world " + std::to_string(world));
	int rc = set_device(d->dev);
	if (rc) return rc;
	// The stream is idle after this read, so the queue can be rearranged.
	if ((rc = read_ctrl(d)))
		return rc;
	// Entries before queue_next have been moved into the particle buffer.
	// Only the unread part [queue_next, q_n) is kept: once the queue is
	// drained it starts again from 0, and growth or compaction copies the
	// unread part alone to the front, so repeated push / run cycles reuse
	// the same memory.
	const uint64_t next = std::min<uint64_t>(d->h_ctrl->queue_next, d->q_n);
	const uint64_t unread = d->q_n - next;
	if (d->q_n + n > d->q_cap || (next > 0 && unread == 0)) {
		const uint64_t need = unread + n;
		uint64_t nc = d->q_cap;
		if (need > d->q_cap)
			nc = std::max<uint64_t>(d->q_cap ? 2 * d->q_cap : 1024, need);
		if (unread > 0 || nc != d->q_cap) {
			// a fresh allocation (the unread part may overlap its new place)
			nbl_particle28* np; uint32_t* nt; uint64_t* ns; uint16_t* nw;
			np = nullptr; nt = nullptr; ns = nullptr; nw = nullptr;
			if ((rc = dev_alloc(&np, nc)) || (rc = dev_alloc(&nt, nc)) || (rc = dev_alloc(&ns, nc)) ||
				(rc = dev_alloc(&nw, nc))) {
				hipFree(np); hipFree(nt); hipFree(ns); hipFree(nw);
				return rc;
			}
			if (unread) {
				HIP_OK(hipMemcpy(np, d->q_part + next, sizeof(nbl_particle28) * unread, hipMemcpyDeviceToDevice));
				HIP_OK(hipMemcpy(nt, d->q_tag + next, sizeof(uint32_t) * unread, hipMemcpyDeviceToDevice));
				HIP_OK(hipMemcpy(ns, d->q_sid + next, sizeof(uint64_t) * unread, hipMemcpyDeviceToDevice));
				HIP_OK(hipMemcpy(nw, d->q_world + next, sizeof(uint16_t) * unread, hipMemcpyDeviceToDevice));
			}
			hipFree(d->q_part); hipFree(d->q_tag); hipFree(d->q_sid); hipFree(d->q_world);
			d->q_part = np; d->q_tag = nt; d->q_sid = ns; d->q_world = nw;
			d->epoch++;
			d->q_cap = nc;
		}
		d->q_n = unread;
		const uint64_t qv[2] = { unread, 0 };   // queue_n, queue_next
		static_assert(offsetof(Ctrl, queue_next) == offsetof(Ctrl, queue_n) + 8, "queue words adjacent");
		HIP_OK(hipMemcpy(&d->ctrl->queue_n, qv, sizeof(qv), hipMemcpyHostToDevice));
		d->h_ctrl->queue_n = unread;
		d->h_ctrl->queue_next = 0;
	}
	// the particles (and given tags) are copied; the running tags, stream
	// ids and world numbers are written on the device
	HIP_OK(hipMemcpy(d->q_part + d->q_n, particles, sizeof(nbl_particle28) * n, hipMemcpyHostToDevice));
	if (tags)
		HIP_OK(hipMemcpy(d->q_tag + d->q_n, tags, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
	if (n) {
		hipLaunchKernelGGL(k_queue_ids, dim3(blocks_for(n)), dim3(256), 0, d->stream, d->q_tag + d->q_n,
			d->q_sid + d->q_n, d->q_world + d->q_n, n, tags ? 0 : 1, d->next_tag, sid_base, (uint16_t)world);
		HIP_OK(hipGetLastError());
		HIP_OK(hipStreamSynchronize(d->stream));
	}
	d->next_tag += n;
	d->q_n += n;
	const uint64_t qn = d->q_n;
	HIP_OK(hipMemcpy(&d->ctrl->queue_n, &qn, sizeof(uint64_t), hipMemcpyHostToDevice));
	if (pushed)
		*pushed = n;
	return NBL_OK;
}

int nbl_driver_push(nbl_driver* d, const nbl_particle28* particles, const uint32_t* tags,
	uint32_t n, uint64_t sid_base, uint32_t* pushed)
{
	return nbl_driver_push_world(d, particles, tags, n, sid_base, 0, pushed);
}

namespace {
// k_step + k_inel in the variant the driver needs (scores, several worlds).
void launch_step_pair(nbl_driver* d, const StepArgs& a, uint32_t g)
{
	const bool multi = d->h_worlds.size() > 1 || d->dep;
	const dim3 grid(g), block(256);
	if (multi && !d->dep && !d->map.np) {
		// several worlds, nothing but records: the launch's particles
		// grouped by world, one world per wave
		const uint32_t nw = (uint32_t)d->h_worlds.size();
		const dim3 ggrid(std::min<uint32_t>(g, 4u * (uint32_t)d->dev->n_cu));
		hipLaunchKernelGGL((k_world_group<false>), ggrid, block, 0, d->stream, d->ctrl, d->buf[0], d->buf[1],
			d->slots, nw, d->wgrp, d->perm);
		hipLaunchKernelGGL(k_world_scan, dim3(1), dim3(256), 0, d->stream, nw, d->wgrp, d->perm);
		if (nw <= NW_LDS)
			hipLaunchKernelGGL((k_world_group<true>), ggrid, block, 0, d->stream, d->ctrl, d->buf[0], d->buf[1],
				d->slots, nw, d->wgrp, d->perm);
		else
			hipLaunchKernelGGL(k_world_place_wide, grid, block, 0, d->stream, d->ctrl, d->buf[0], d->buf[1],
				d->slots, nw, d->wgrp, d->perm);
		hipLaunchKernelGGL((k_step<false, false, true>), grid, block, 0, d->stream, a);
		hipLaunchKernelGGL((k_inel<false, true>), grid, block, 0, d->stream, a);
	} else if (d->map.np && multi) {
		hipLaunchKernelGGL((k_step<true, true>), grid, block, 0, d->stream, a);
		hipLaunchKernelGGL((k_inel<true, true>), grid, block, 0, d->stream, a);
	} else if (d->map.np) {
		hipLaunchKernelGGL((k_step<true, false>), grid, block, 0, d->stream, a);
		hipLaunchKernelGGL((k_inel<true, false>), grid, block, 0, d->stream, a);
	} else if (multi) {
		hipLaunchKernelGGL((k_step<false, true>), grid, block, 0, d->stream, a);
		hipLaunchKernelGGL((k_inel<false, true>), grid, block, 0, d->stream, a);
	} else {
		hipLaunchKernelGGL((k_step<false, false>), grid, block, 0, d->stream, a);
		hipLaunchKernelGGL((k_inel<false, false>), grid, block, 0, d->stream, a);
	}
}
} // namespace

namespace {
StepArgs step_args(const nbl_driver* d)
{
	StepArgs a;
	a.W = d->W;
	a.buf[0] = d->buf[0];
	a.buf[1] = d->buf[1];
	a.spill = d->spill;
	a.spill_cap = d->spill_cap;
	a.inel = d->inel;
	a.ctrl = d->ctrl;
	a.worlds = d->d_worlds;
	a.perm = d->perm;
	a.perm_n = d->wgrp ? d->wgrp + 3 * d->h_worlds.size() : nullptr;
	a.det = d->det;
	a.det_sc = d->det_sc;
	a.dep = d->dep;
	a.dep_cap = (uint32_t)std::min<uint64_t>(d->dep_cap, 0xffffffffu);
	a.map = d->map;
	a.det_cap = d->det_cap;
	a.slots = d->slots;
	a.K = d->K;
	return a;
}

// One tracking iteration: control, refill, step pair.  The step pair is
// bracketed by timing events when timing is on.
int enqueue_iteration(nbl_driver* d, const StepArgs& a)
{
	const uint32_t g_refill = blocks_for(d->capacity);
	// grid-stride: blocks_per_cu x compute units blocks of 256 threads (a
	// multiple of the resident blocks), never more than the capacity needs
	const uint32_t g_step = std::min<uint32_t>(blocks_for(d->capacity), d->blocks_per_cu * (uint32_t)d->dev->n_cu);
	hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(64), 0, d->stream, d->ctrl, d->slots, d->spill_cap);
	hipLaunchKernelGGL(k_refill, dim3(g_refill), dim3(256), 0, d->stream, d->ctrl, d->buf[0],
		d->buf[1], d->spill, d->q_part, d->q_tag, d->q_sid, d->q_world, (uint32_t)d->map.np, 0ull);
	std::pair<hipEvent_t, hipEvent_t> ev{};
	if (d->timing) {
		if (d->ev_free.empty()) {
			HIP_OK(hipEventCreate(&ev.first));
			HIP_OK(hipEventCreate(&ev.second));
		} else {
			ev = d->ev_free.back();
			d->ev_free.pop_back();
		}
		HIP_OK(hipEventRecord(ev.first, d->stream));
	}
	launch_step_pair(d, a, g_step);
	if (d->timing) {
		HIP_OK(hipEventRecord(ev.second, d->stream));
		d->ev_used.push_back(ev);
		d->step_launches++;
	}
	return NBL_OK;
}

// n iterations as one captured graph (built on first use for this n and
// this set of launch arguments).  The iterations contain kernels only.
int launch_graph(nbl_driver* d, const StepArgs& a, uint32_t n)
{
	for (auto& gr : d->graphs)
		if (gr.n == n && gr.epoch == d->epoch) {
			HIP_OK(hipGraphLaunch(gr.exec, d->stream));
			return NBL_OK;
		}
	// drop graphs made with stale arguments
	for (auto it = d->graphs.begin(); it != d->graphs.end();) {
		if (it->epoch != d->epoch) {
			hipGraphExecDestroy(it->exec);
			hipGraphDestroy(it->g);
			it = d->graphs.erase(it);
		} else {
			++it;
		}
	}
	nbl_driver::Graph gr{ n, d->epoch, nullptr, nullptr };
	HIP_OK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
	int rc = NBL_OK;
	for (uint32_t it = 0; it < n && rc == NBL_OK; ++it)
		rc = enqueue_iteration(d, a);
	const hipError_t e = hipStreamEndCapture(d->stream, &gr.g);
	if (rc) return rc;
	if (e != hipSuccess)
		return fail(NBL_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
	HIP_OK(hipGraphInstantiate(&gr.exec, gr.g, nullptr, nullptr, 0));
	d->graphs.push_back(gr);
	HIP_OK(hipGraphLaunch(gr.exec, d->stream));
	return NBL_OK;
}
} // namespace

int nbl_driver_iterate(nbl_driver* d, uint32_t n_iter)
{
	clear_error();
	if (!d)
		return fail(NBL_ERR_ARG, "null driver");
	int rc = set_device(d->dev);
	if (rc) return rc;
	// an error left by an earlier, unrelated runtime call is not this call's
	(void)hipGetLastError();
	const StepArgs a = step_args(d);
	if (d->use_graph && !d->timing && n_iter > 1) {
		if ((rc = launch_graph(d, a, n_iter)))
			return rc;
	} else {
		for (uint32_t it = 0; it < n_iter; ++it)
			if ((rc = enqueue_iteration(d, a)))
				return rc;
	}
	HIP_OK(hipGetLastError());
	return NBL_OK;
}

// a candidate launch configuration must beat the incumbent by this much
constexpr double NBL_TUNE_MARGIN = 0.03;

int nbl_driver_autotune(nbl_driver* d, uint32_t iters_per_trial, nbl_tune_result* out)
{
	clear_error();
	if (!d || !out || iters_per_trial < 1)
		return fail(NBL_ERR_ARG, "invalid autotune arguments");
	int rc = set_device(d->dev);
	if (rc) return rc;
	if ((rc = read_ctrl(d)))
		return rc;
	if (d->h_ctrl->queue_n == d->h_ctrl->queue_next)
		return fail(NBL_ERR_ARG, "autotune measures queued primaries: push the work first");
	const bool graph = d->use_graph, timing = d->timing;
	const double step_ms0 = d->step_ms;
	const uint64_t launches0 = d->step_launches;
	d->use_graph = false;
	d->timing = true;
	// the caller's timing totals do not include the trials
	auto restore = [&]() {
		d->use_graph = graph;
		d->timing = timing;
		d->step_ms = step_ms0;
		d->step_launches = launches0;
	};
	struct Cfg { uint32_t K, bpc, target; };
	Cfg best{ d->K, d->blocks_per_cu, d->target };
	double best_rate = -1.0;
	uint32_t trials = 0;
	bool stop = false;
	// particle-steps per second of step-kernel time (HIP events around each
	// k_step + k_inel pair) over `iters_per_trial` iterations of the queued
	// work; host round trips of the short trials would otherwise favour the
	// configurations that do more work per launch
	auto trial = [&](const Cfg& c, double& rate) -> int {
		int r = nbl_driver_set_launch(d, c.K, c.bpc, c.target);
		if (r) return r;
		if ((r = read_ctrl(d))) return r;
		const uint64_t s0 = total_steps(*d->h_ctrl);
		const double ms0 = d->step_ms;
		if ((r = nbl_driver_iterate(d, iters_per_trial))) return r;
		if ((r = read_ctrl(d))) return r;
		const double dt = (d->step_ms - ms0) * 1e-3;
		rate = dt > 0 ? (double)(total_steps(*d->h_ctrl) - s0) / dt : 0.0;
		++trials;
		// stop before the detected buffer or the queue runs out
		const uint64_t left = d->h_ctrl->queue_n - d->h_ctrl->queue_next;
		if (d->h_ctrl->det_ctr > d->det_cap / 2 || left == 0)
			stop = true;
		return NBL_OK;
	};
	// ramp the population up with the current configuration
	double rate = 0;
	if ((rc = trial(best, rate))) { restore(); return rc; }
	if ((rc = trial(best, rate))) { restore(); return rc; }
	best_rate = rate;
	const uint32_t Ks[] = { 4, 8, 16, 32 };
	const uint32_t Bs[] = { 8, 16, 32, 64 };
	const uint32_t Ts[] = { std::max<uint32_t>(64, d->capacity / 4), std::max<uint32_t>(64, d->capacity / 2),
		d->capacity };
	// Coordinate search (steps per launch, then grid, then particles in
	// flight).  The work mix drifts as the run goes on, so each candidate is
	// measured right next to the incumbent and replaces it only when faster
	// by more than NBL_TUNE_MARGIN.
	for (int dim = 0; dim < 3 && !stop; ++dim) {
		const int nc = dim == 2 ? 3 : 4;
		for (int i = 0; i < nc && !stop; ++i) {
			Cfg c = best;
			if (dim == 0) c.K = Ks[i]; else if (dim == 1) c.bpc = Bs[i]; else c.target = Ts[i];
			if (c.K == best.K && c.bpc == best.bpc && c.target == best.target)
				continue;
			double rc_c = 0, rc_b = 0;
			if ((rc = trial(c, rc_c))) { restore(); return rc; }
			if ((rc = trial(best, rc_b))) { restore(); return rc; }
			best_rate = rc_b;
			if (rc_c > rc_b * (1.0 + NBL_TUNE_MARGIN)) {
				best = c;
				best_rate = rc_c;
			}
		}
	}
	restore();
	if ((rc = nbl_driver_set_launch(d, best.K, best.bpc, best.target)))
		return rc;
	out->steps_per_launch = best.K;
	out->blocks_per_cu = best.bpc;
	out->target = best.target;
	out->trials = trials;
	out->steps_per_s = best_rate;
	return NBL_OK;
}

int nbl_driver_set_deposits(nbl_driver* d, int enable)
{
	clear_error();
	if (!d)
		return fail(NBL_ERR_ARG, "null driver");
	int rc = set_device(d->dev);
	if (rc) return rc;
	HIP_OK(hipStreamSynchronize(d->stream));
	hipFree(d->dep);
	d->dep = nullptr;
	d->dep_cap = 0;
	d->dep_per_particle = 0;
	d->epoch++;
	if (!enable)
		return NBL_OK;
	// one record per step and one per inelastic event, for every particle
	// the buffer can hold
	d->dep_per_particle = d->K + 1;
	d->dep_cap = (uint64_t)d->slots * d->dep_per_particle;
	if (d->dep_cap > 0xffffffffull)
		return fail(NBL_ERR_CAPACITY, "deposit buffer too large for this capacity and steps per launch");
	if ((rc = dev_alloc(&d->dep, d->dep_cap)))
		return rc;
	const uint32_t zero = 0;
	HIP_OK(hipMemcpy(&d->ctrl->dep_ctr, &zero, sizeof(zero), hipMemcpyHostToDevice));
	return NBL_OK;
}

int nbl_driver_drain_deposits(nbl_driver* d, nbl_deposit* out, uint64_t cap, uint64_t* n)
{
	clear_error();
	if (!d || !n || (cap && !out))
		return fail(NBL_ERR_ARG, "invalid drain arguments");
	if (!d->dep)
		return fail(NBL_ERR_ARG, "deposits are not enabled on this driver");
	int rc = set_device(d->dev);
	if (rc) return rc;
	if ((rc = read_ctrl(d)))
		return rc;
	const uint64_t written = d->h_ctrl->dep_ctr;
	if (written > d->dep_cap)
		return fail(NBL_ERR_CAPACITY, "energy-deposit records were dropped: drain after every iteration");
	if (written > cap) {
		*n = written;
		return fail(NBL_ERR_CAPACITY, "output array too small for " + std::to_string(written) + " records");
	}
	if (written)
		HIP_OK(hipMemcpyAsync(out, d->dep, sizeof(nbl_deposit) * written, hipMemcpyDeviceToHost, d->stream));
	const uint32_t zero = 0;
	HIP_OK(hipMemcpyAsync(&d->ctrl->dep_ctr, &zero, sizeof(zero), hipMemcpyHostToDevice, d->stream));
	HIP_OK(hipStreamSynchronize(d->stream));
	*n = written;
	return NBL_OK;
}

int nbl_driver_set_graph(nbl_driver* d, int enable)
{
	clear_error();
	if (!d)
		return fail(NBL_ERR_ARG, "null driver");
	d->use_graph = enable != 0;
	return NBL_OK;
}

int nbl_driver_set_launch(nbl_driver* d, uint32_t steps_per_launch, uint32_t blocks_per_cu, uint32_t target)
{
	clear_error();
	if (!d)
		return fail(NBL_ERR_ARG, "null driver");
	if (steps_per_launch < 1 || steps_per_launch > 1024)
		return fail(NBL_ERR_ARG, "steps_per_launch must be in [1, 1024]");
	if (blocks_per_cu < 1 || blocks_per_cu > 128)
		return fail(NBL_ERR_ARG, "blocks_per_cu must be in [1, 128]");
	if (target < 1 || target > d->capacity)
		return fail(NBL_ERR_ARG, "target must be in [1, capacity]");
	if (d->dep && steps_per_launch + 1 > d->dep_per_particle)
		return fail(NBL_ERR_ARG, "deposits are enabled for fewer steps per launch: re-enable them after this change");
	int rc = set_device(d->dev);
	if (rc) return rc;
	if (d->K != steps_per_launch || d->blocks_per_cu != blocks_per_cu)
		d->epoch++;
	d->K = steps_per_launch;
	d->blocks_per_cu = blocks_per_cu;
	// the control block is only read by kernels: update it between launches
	HIP_OK(hipStreamSynchronize(d->stream));
	HIP_OK(hipMemcpy(&d->ctrl->target, &target, sizeof(uint32_t), hipMemcpyHostToDevice));
	d->target = target;
	return NBL_OK;
}

namespace {
// The spill stack holds secondaries that found no room in the particle
// buffer.  Deep cascades (keV primaries make hundreds of secondaries each)
// can fill it, and a particle whose secondary finds no room anywhere waits:
// with the buffer and the stack full of such particles nothing would move.
// Whenever the host reads the counters (stream idle), a stack more than
// half full is doubled, so secondaries always find room.
int grow_spill(nbl_driver* d)
{
	const Ctrl& c = *d->h_ctrl;
	const uint32_t used = c.stepped ? std::min(c.spill_ctr, d->spill_cap) : c.spill_n;
	if (used <= d->spill_cap / 2 || d->spill_cap >= (1u << 31))
		return NBL_OK;
	const uint32_t ncap = d->spill_cap * 2u;
	Bufs nb;
	void* nmem = nullptr;
	int rc = alloc_bufs(nb, nmem, ncap);
	if (rc) return rc;
	double* nsc = nullptr;
	if (d->map.np && (rc = dev_alloc(&nsc, (size_t)d->map.np * ncap))) {
		hipFree(nmem);
		return rc;
	}
	const size_t on = d->spill_cap, nn = ncap, u = used;
	char* ob = d->spill.base;
	char* nbb = nb.base;
	if (u) {
		HIP_OK(hipMemcpyAsync(nbb, ob, 8 * u, hipMemcpyDeviceToDevice, d->stream));
		for (int f = 0; f < N_FIELDS; ++f)
			HIP_OK(hipMemcpyAsync(nbb + 8 * nn + 4 * f * nn, ob + 8 * on + 4 * f * on, 4 * u,
				hipMemcpyDeviceToDevice, d->stream));
		for (int q = 0; q < d->map.np; ++q)
			HIP_OK(hipMemcpyAsync(nsc + q * nn, d->spill.sc + q * on, 8 * u, hipMemcpyDeviceToDevice,
				d->stream));
	}
	// secondaries beyond the old top were never stored (their parents wait
	// and retry): the counter must not count them once the stack is larger
	if (c.stepped && c.spill_ctr > used) {
		const uint32_t top = used;
		HIP_OK(hipMemcpyAsync(&d->ctrl->spill_ctr, &top, sizeof(top), hipMemcpyHostToDevice, d->stream));
		d->h_ctrl->spill_ctr = top;
	}
	HIP_OK(hipStreamSynchronize(d->stream));
	hipFree(d->spill_mem);
	if (d->map.np)
		hipFree(d->sc_mem[2]);
	d->spill = nb;
	d->spill_mem = nmem;
	if (d->map.np) {
		d->spill.sc = nsc;
		d->sc_mem[2] = nsc;
	}
	d->spill_cap = ncap;
	d->spill_growths++;
	d->epoch++;
	return NBL_OK;
}
} // namespace

int nbl_driver_counts(nbl_driver* d, uint32_t* running, uint32_t* detected, uint64_t* queued)
{
	clear_error();
	if (!d)
		return fail(NBL_ERR_ARG, "null driver");
	int rc = set_device(d->dev);
	if (rc) return rc;
	if ((rc = read_ctrl(d)))
		return rc;
	if ((rc = grow_spill(d)))
		return rc;
	const Ctrl& c = *d->h_ctrl;
	if (running) *running = active_count(c, d->spill_cap);
	if (detected) *detected = std::min(c.det_ctr, d->det_cap);
	if (queued) *queued = c.queue_n - c.queue_next;
	return NBL_OK;
}

namespace {
// Copy out the detected records (and their scores, if `scores`) and clear
// the device buffer.
int drain(nbl_driver* d, nbl_detected* out, double* scores, uint32_t cap, uint32_t* n)
{
	int rc = read_ctrl(d);
	if (rc)
		return rc;
	const uint32_t have = std::min(d->h_ctrl->det_ctr, d->det_cap);
	if (have > cap) {
		*n = have;
		return fail(NBL_ERR_CAPACITY, "output array too small for " + std::to_string(have) + " records");
	}
	if (have) {
		HIP_OK(hipMemcpyAsync(out, d->det, sizeof(nbl_detected) * have, hipMemcpyDeviceToHost, d->stream));
		if (scores && d->map.np)
			HIP_OK(hipMemcpyAsync(scores, d->det_sc, sizeof(double) * d->map.np * (size_t)have,
				hipMemcpyDeviceToHost, d->stream));
	}
	hipLaunchKernelGGL(k_reset_det, dim3(1), dim3(64), 0, d->stream, d->ctrl);
	HIP_OK(hipStreamSynchronize(d->stream));
	*n = have;
	return NBL_OK;
}
} // namespace

int nbl_driver_drain_detected(nbl_driver* d, nbl_detected* out, uint32_t cap, uint32_t* n)
{
	clear_error();
	if (!d || !n || (cap && !out))
		return fail(NBL_ERR_ARG, "invalid drain arguments");
	int rc = set_device(d->dev);
	if (rc) return rc;
	return drain(d, out, nullptr, cap, n);
}

int nbl_driver_drain_detected_scores(nbl_driver* d, nbl_detected* out, double* scores, uint32_t cap,
	uint32_t* n)
{
	clear_error();
	if (!d || !n || (cap && (!out || !scores)))
		return fail(NBL_ERR_ARG, "invalid drain arguments");
	if (!d->map.np)
		return fail(NBL_ERR_ARG, "scores are not enabled on this driver");
	int rc = set_device(d->dev);
	if (rc) return rc;
	return drain(d, out, scores, cap, n);
}

int nbl_driver_steps(nbl_driver* d, uint64_t* steps)
{
	clear_error();
	if (!d || !steps)
		return fail(NBL_ERR_ARG, "null argument");
	int rc = set_device(d->dev);
	if (rc) return rc;
	if ((rc = read_ctrl(d)))
		return rc;
	*steps = total_steps(*d->h_ctrl);
	return NBL_OK;
}

namespace {
// The frame loop shared by nbl_driver_run and nbl_driver_run_frames: track
// everything pushed to the end.  Each frame's detected records (and
// scores) are copied on the stream into pinned staging and the record
// counter reset, the next frame is enqueued at once, and the host hands the
// staged records to `sink` while the GPU runs that frame.  Frames are
// `frame_iters` iterations, or (0) 4 growing to 32.
using RecordSink = std::function<int(const nbl_detected*, const double*, uint32_t, uint32_t, uint64_t)>;
int run_loop(nbl_driver* d, bool with_scores, uint32_t frame_iters, const RecordSink& sink)
{
	const size_t np = with_scores ? (size_t)d->map.np : 0;
	if (!d->h_stage &&
		hipHostMalloc((void**)&d->h_stage, sizeof(nbl_detected) * d->det_cap, hipHostMallocDefault) != hipSuccess) {
		d->h_stage = nullptr;
		return fail(NBL_ERR_ALLOC, "pinned staging for detected records");
	}
	if (np && d->stage_sc_n < np * d->det_cap) {
		hipHostFree(d->h_stage_sc);
		d->stage_sc_n = 0;
		if (hipHostMalloc((void**)&d->h_stage_sc, sizeof(double) * np * d->det_cap, hipHostMallocDefault)
			!= hipSuccess) {
			d->h_stage_sc = nullptr;
			return fail(NBL_ERR_ALLOC, "pinned staging for detected scores");
		}
		d->stage_sc_n = np * d->det_cap;
	}
	if (!d->stage_ev)
		HIP_OK(hipEventCreateWithFlags(&d->stage_ev, hipEventDisableTiming));
	uint32_t frame = frame_iters ? frame_iters : 4;
	uint64_t last_steps = ~0ull;
	int stalled = 0;
	int rc = nbl_driver_iterate(d, frame);
	if (rc) return rc;
	for (;;) {
		uint32_t running = 0, det = 0;
		uint64_t queued = 0;
		if ((rc = nbl_driver_counts(d, &running, &det, &queued)))
			return rc;
		if (det) {
			HIP_OK(hipMemcpyAsync(d->h_stage, d->det, sizeof(nbl_detected) * det, hipMemcpyDeviceToHost,
				d->stream));
			if (np)
				HIP_OK(hipMemcpyAsync(d->h_stage_sc, d->det_sc, sizeof(double) * np * (size_t)det,
					hipMemcpyDeviceToHost, d->stream));
			hipLaunchKernelGGL(k_reset_det, dim3(1), dim3(64), 0, d->stream, d->ctrl);
			HIP_OK(hipEventRecord(d->stage_ev, d->stream));
		}
		// particles waiting for a record slot still count as running
		const bool finished = running == 0 && queued == 0;
		if (!finished) {
			// Every particle waiting for a secondary slot with the buffer and
			// the spill stack full would never finish: report it instead of
			// spinning.
			if (total_steps(*d->h_ctrl) == last_steps && det == 0) {
				if (++stalled >= 4)
					return fail(NBL_ERR_CAPACITY, "no progress: particle buffer and spill stack full of "
						"particles waiting for secondary-electron slots; raise the capacity");
			} else {
				stalled = 0;
			}
			last_steps = total_steps(*d->h_ctrl);
			if (!frame_iters)
				frame = std::min<uint32_t>(frame * 2, 32);
			if ((rc = nbl_driver_iterate(d, frame)))
				return rc;
		}
		if (det)
			HIP_OK(hipEventSynchronize(d->stage_ev));
		if ((rc = sink(d->h_stage, np ? d->h_stage_sc : nullptr, det, running, queued)))
			return rc;
		if (finished)
			break;
	}
	HIP_OK(hipStreamSynchronize(d->stream));
	return NBL_OK;
}

int run(nbl_driver* d, nbl_detected* out, double* scores, uint64_t cap, uint64_t* n_detected)
{
	const size_t np = scores ? (size_t)d->map.np : 0;
	uint64_t total = 0;
	const int rc = run_loop(d, scores != nullptr, 0,
		[&](const nbl_detected* recs, const double* sc, uint32_t n, uint32_t, uint64_t) -> int {
			const uint64_t room = cap > total ? std::min<uint64_t>(cap - total, n) : 0;
			if (room) {
				std::memcpy(out + total, recs, sizeof(nbl_detected) * room);
				if (np)
					std::memcpy(scores + total * np, sc, sizeof(double) * np * room);
			}
			total += n;
			return NBL_OK;
		});
	*n_detected = total;
	if (rc)
		return rc;
	if (total > cap)
		return fail(NBL_ERR_CAPACITY, "more detected electrons than the output array holds");
	return NBL_OK;
}
} // namespace

int nbl_driver_run(nbl_driver* d, nbl_detected* out, uint64_t cap, uint64_t* n_detected)
{
	clear_error();
	if (!d || !n_detected || (cap && !out))
		return fail(NBL_ERR_ARG, "invalid run arguments");
	return run(d, out, nullptr, cap, n_detected);
}

int nbl_driver_run_scores(nbl_driver* d, nbl_detected* out, double* scores, uint64_t cap,
	uint64_t* n_detected)
{
	clear_error();
	if (!d || !n_detected || (cap && (!out || !scores)))
		return fail(NBL_ERR_ARG, "invalid run arguments");
	if (!d->map.np)
		return fail(NBL_ERR_ARG, "scores are not enabled on this driver");
	return run(d, out, scores, cap, n_detected);
}

int nbl_driver_run_frames(nbl_driver* d, uint32_t frame_iters, nbl_record_sink sink, void* user)
{
	clear_error();
	if (!d || !sink || frame_iters < 1 || frame_iters > 1u << 16)
		return fail(NBL_ERR_ARG, "invalid run_frames arguments (frame_iters in [1, 65536], a sink)");
	int rc = set_device(d->dev);
	if (rc) return rc;
	return run_loop(d, false, frame_iters,
		[&](const nbl_detected* recs, const double*, uint32_t n, uint32_t running, uint64_t queued) -> int {
			if (sink(user, recs, n, running, queued) != 0)
				return fail(NBL_ERR_ARG, "the record sink stopped the run");
			return NBL_OK;
		});
}

int nbl_driver_set_material_params(nbl_driver* d, int32_t material, float el_mfp_scale,
	float in_mfp_scale, float barrier)
{
	clear_error();
	if (!d || material < 0 || material >= (int32_t)d->W.nmats)
		return fail(NBL_ERR_ARG, "no such material");
	if (!(el_mfp_scale > 0.0f) || !(in_mfp_scale > 0.0f) || !std::isfinite(el_mfp_scale) ||
		!std::isfinite(in_mfp_scale) || !std::isfinite(barrier))
		return fail(NBL_ERR_ARG, "mean-free-path scales must be positive and finite, the barrier finite");
	int rc = set_device(d->dev);
	if (rc) return rc;
	// the material arrays are read by kernels in flight: change them when idle
	HIP_OK(hipStreamSynchronize(d->stream));
	auto apply = [&](Material& m) {
		m.el_scale = el_mfp_scale;
		m.in_scale = in_mfp_scale;
		m.barrier = barrier;
	};
	apply(d->h_mats[(size_t)material]);
	HIP_OK(hipMemcpy(d->d_mats + material, &d->h_mats[(size_t)material], sizeof(Material), hipMemcpyHostToDevice));
	// the same material index of every other world (the scans of a batch
	// share their material list)
	for (size_t w = 0; w < d->world_mats.size(); ++w) {
		std::vector<Material>& hm = d->h_world_mats[w];
		if ((size_t)material >= hm.size() || (int)w + 1 >= (int)d->h_worlds.size() ||
			material >= d->h_worlds[w + 1].nmats)
			continue;
		apply(hm[(size_t)material]);
		HIP_OK(hipMemcpy(d->world_mats[w] + material, &hm[(size_t)material], sizeof(Material), hipMemcpyHostToDevice));
	}
	return NBL_OK;
}

int nbl_driver_get_material_params(nbl_driver* d, int32_t material, float* el_mfp_scale,
	float* in_mfp_scale, float* barrier)
{
	clear_error();
	if (!d || material < 0 || material >= (int32_t)d->W.nmats || !el_mfp_scale || !in_mfp_scale || !barrier)
		return fail(NBL_ERR_ARG, "invalid arguments");
	const Material& m = d->h_mats[(size_t)material];
	*el_mfp_scale = m.el_scale;
	*in_mfp_scale = m.in_scale;
	*barrier = m.barrier;
	return NBL_OK;
}

int nbl_driver_set_score_params(nbl_driver* d, const int32_t* materials, const int32_t* kinds, int32_t n)
{
	clear_error();
	if (!d || n < 0 || n > MAX_PARAMS || (n && (!materials || !kinds)))
		return fail(NBL_ERR_ARG, "between 0 and 8 parameters");
	ScoreMap map;
	map.np = n;
	std::memset(map.pidx, -1, sizeof(map.pidx));
	for (int32_t i = 0; i < n; ++i) {
		if (materials[i] < 0 || materials[i] >= std::min<int32_t>(d->W.nmats, MAX_SCORED_MATERIALS))
			return fail(NBL_ERR_ARG, "parameter " + std::to_string(i) + ": no such material (at most 16 carry parameters)");
		if (kinds[i] < 0 || kinds[i] > 2)
			return fail(NBL_ERR_ARG, "parameter " + std::to_string(i) + ": kind must be 0, 1 or 2");
		if (map.pidx[materials[i]][kinds[i]] >= 0)
			return fail(NBL_ERR_ARG, "parameter " + std::to_string(i) + " repeats an earlier one");
		map.pidx[materials[i]][kinds[i]] = (int8_t)i;
	}
	int rc = set_device(d->dev);
	if (rc) return rc;
	if ((rc = read_ctrl(d)))
		return rc;
	if (active_count(*d->h_ctrl, d->spill_cap) != 0 || d->h_ctrl->queue_n != d->h_ctrl->queue_next ||
		d->h_ctrl->det_ctr != 0)
		return fail(NBL_ERR_ARG, "scores can only be switched while no particle is in flight or detected");
	for (double*& p : d->sc_mem) { hipFree(p); p = nullptr; }
	hipFree(d->det_sc);
	d->det_sc = nullptr;
	Bufs* bufs[4] = { &d->buf[0], &d->buf[1], &d->spill, &d->inel };
	for (Bufs* b : bufs)
		b->sc = nullptr;
	d->map.np = 0;
	d->epoch++;
	if (n == 0)
		return NBL_OK;
	for (int k = 0; k < 4; ++k) {
		if ((rc = dev_alloc(&d->sc_mem[k], (size_t)n * bufs[k]->sc_stride)))
			return rc;
		bufs[k]->sc = d->sc_mem[k];
	}
	if ((rc = dev_alloc(&d->det_sc, (size_t)n * d->det_cap)))
		return rc;
	d->map = map;
	return NBL_OK;
}

void* nbl_driver_stream(nbl_driver* d) { return d ? (void*)d->stream : nullptr; }

int nbl_driver_get_info(nbl_driver* d, nbl_driver_info* out)
{
	clear_error();
	if (!d || !out)
		return fail(NBL_ERR_ARG, "null argument");
	out->queue_capacity = d->q_cap;
	out->queue_entries = d->q_n;
	out->slots = d->slots;
	out->spill_capacity = d->spill_cap;
	out->spill_growths = d->spill_growths;
	out->detected_capacity = d->det_cap;
	return NBL_OK;
}

int nbl_geometry_query(const nbl_geometry* g, const float* pos, const float* dir, const float* dist,
	const int32_t* ignore_tri, const int32_t* ignore_mat, int32_t* out_tri, float* out_dist, uint32_t n)
{
	clear_error();
	if (!g || (n && (!pos || !dir || !dist || !ignore_tri || !ignore_mat || !out_tri || !out_dist)))
		return fail(NBL_ERR_ARG, "invalid query arguments");
	if (!n)
		return NBL_OK;
	int rc = set_device(g->dev);
	if (rc) return rc;
	for (uint32_t i = 0; i < n; ++i)
		if (ignore_tri[i] < -1 || ignore_tri[i] >= g->g.ntri)
			return fail(NBL_ERR_ARG, "ignore_tri out of range");
	float *dp, *dd, *ddist, *dout;
	int *dit, *dim, *dtri;
	if ((rc = dev_alloc(&dp, 3 * (size_t)n)) || (rc = dev_alloc(&dd, 3 * (size_t)n)) ||
		(rc = dev_alloc(&ddist, n)) || (rc = dev_alloc(&dit, n)) || (rc = dev_alloc(&dim, n)) ||
		(rc = dev_alloc(&dtri, n)) || (rc = dev_alloc(&dout, n)))
		return rc;
	hipStream_t s = g->dev->stream;
	HIP_OK(hipMemcpyAsync(dp, pos, 12 * (size_t)n, hipMemcpyHostToDevice, s));
	HIP_OK(hipMemcpyAsync(dd, dir, 12 * (size_t)n, hipMemcpyHostToDevice, s));
	HIP_OK(hipMemcpyAsync(ddist, dist, 4 * (size_t)n, hipMemcpyHostToDevice, s));
	HIP_OK(hipMemcpyAsync(dit, ignore_tri, 4 * (size_t)n, hipMemcpyHostToDevice, s));
	HIP_OK(hipMemcpyAsync(dim, ignore_mat, 4 * (size_t)n, hipMemcpyHostToDevice, s));
	(void)hipGetLastError();   // not this call's
	hipLaunchKernelGGL(k_query, dim3(blocks_for(n)), dim3(256), 0, s, g->g, dp, dd, ddist, dit, dim,
		dtri, dout, n);
	HIP_OK(hipGetLastError());
	HIP_OK(hipMemcpyAsync(out_tri, dtri, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipMemcpyAsync(out_dist, dout, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipStreamSynchronize(s));
	hipFree(dp); hipFree(dd); hipFree(ddist); hipFree(dit); hipFree(dim); hipFree(dtri); hipFree(dout);
	return NBL_OK;
}

int nbl_table_probe(const nbl_material* m, int32_t which, int32_t mode, const float* x,
	const float* y, const float* z, float* out, uint32_t n)
{
	clear_error();
	if (!m || which < 0 || which > 6 || mode < 0 || mode > 2 || (n && (!x || !out)))
		return fail(NBL_ERR_ARG, "invalid probe arguments");
	if (!m->tables[which])
		return fail(NBL_ERR_ARG, "material has no such table");
	if (!n)
		return NBL_OK;
	const bool two = (which == 1 || which == 3 || which == 6);
	const bool three = (which == 4 || which == 5);
	if ((two || three) && !y)
		return fail(NBL_ERR_ARG, "y needed");
	if (three && !z)
		return fail(NBL_ERR_ARG, "z needed");
	if (mode == 2 && !two)
		return fail(NBL_ERR_ARG, "get_nearest exists for 2-D tables only");
	int rc = set_device(m->dev);
	if (rc) return rc;
	float *dx, *dy, *dz, *dout;
	if ((rc = dev_alloc(&dx, n)) || (rc = dev_alloc(&dy, n)) || (rc = dev_alloc(&dz, n)) ||
		(rc = dev_alloc(&dout, n)))
		return rc;
	hipStream_t s = m->dev->stream;
	HIP_OK(hipMemcpyAsync(dx, x, 4 * (size_t)n, hipMemcpyHostToDevice, s));
	HIP_OK(hipMemcpyAsync(dy, y ? y : x, 4 * (size_t)n, hipMemcpyHostToDevice, s));
	HIP_OK(hipMemcpyAsync(dz, z ? z : x, 4 * (size_t)n, hipMemcpyHostToDevice, s));
	(void)hipGetLastError();   // not this call's
	hipLaunchKernelGGL(k_probe, dim3(blocks_for(n)), dim3(256), 0, s, m->m, which, mode, dx, dy, dz, dout, n);
	HIP_OK(hipGetLastError());
	HIP_OK(hipMemcpyAsync(out, dout, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipStreamSynchronize(s));
	hipFree(dx); hipFree(dy); hipFree(dz); hipFree(dout);
	return NBL_OK;
}

int nbl_driver_step_states(nbl_driver* d, nbl_step_state* states, int32_t* event, int32_t* outcome,
	nbl_step_state* se, int32_t* se_made, uint32_t n)
{
	clear_error();
	if (!d || (n && (!states || !event || !outcome || !se || !se_made)))
		return fail(NBL_ERR_ARG, "invalid step arguments");
	if (!n)
		return NBL_OK;
	for (uint32_t i = 0; i < n; ++i) {
		if (states[i].mat >= d->W.nmats || (states[i].mat < 0 && states[i].mat < -128))
			return fail(NBL_ERR_ARG, "state material out of range");
		if (states[i].last_tri < -1 || states[i].last_tri >= d->W.g.ntri)
			return fail(NBL_ERR_ARG, "state last_tri out of range");
	}
	int rc = set_device(d->dev);
	if (rc) return rc;
	nbl_step_state *dst, *dse;
	int *dev_ev, *dev_oc, *dev_made;
	if ((rc = dev_alloc(&dst, n)) || (rc = dev_alloc(&dse, n)) || (rc = dev_alloc(&dev_ev, n)) ||
		(rc = dev_alloc(&dev_oc, n)) || (rc = dev_alloc(&dev_made, n)))
		return rc;
	hipStream_t s = d->stream;
	HIP_OK(hipMemcpyAsync(dst, states, sizeof(nbl_step_state) * n, hipMemcpyHostToDevice, s));
	(void)hipGetLastError();   // not this call's
	hipLaunchKernelGGL(k_step_states, dim3(blocks_for(n)), dim3(256), 0, s, d->W, dst, dev_ev, dev_oc,
		dse, dev_made, n);
	HIP_OK(hipGetLastError());
	HIP_OK(hipMemcpyAsync(states, dst, sizeof(nbl_step_state) * n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipMemcpyAsync(se, dse, sizeof(nbl_step_state) * n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipMemcpyAsync(event, dev_ev, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipMemcpyAsync(outcome, dev_oc, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipMemcpyAsync(se_made, dev_made, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
	HIP_OK(hipStreamSynchronize(s));
	hipFree(dst); hipFree(dse); hipFree(dev_ev); hipFree(dev_oc); hipFree(dev_made);
	return NBL_OK;
}

int nbl_host_alloc(nbl_device* dev, uint64_t bytes, void** ptr)
{
	clear_error();
	if (!dev || !ptr)
		return fail(NBL_ERR_ARG, "invalid host allocation arguments");
	*ptr = nullptr;
	if (!bytes)
		return NBL_OK;
	int rc = set_device(dev);
	if (rc) return rc;
	void* p = nullptr;
	if (hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) != hipSuccess || !p) {
		(void)hipGetLastError();
		return fail(NBL_ERR_ALLOC, "page-locked host memory of " + std::to_string(bytes) + " bytes");
	}
	*ptr = p;
	return NBL_OK;
}

int nbl_host_free(void* ptr)
{
	clear_error();
	if (ptr)
		HIP_OK(hipHostFree(ptr));
	return NBL_OK;
}

int nbl_device_memory(nbl_device* dev, uint64_t* free_bytes, uint64_t* total_bytes)
{
	clear_error();
	if (!dev || !free_bytes || !total_bytes)
		return fail(NBL_ERR_ARG, "invalid memory query arguments");
	int rc = set_device(dev);
	if (rc) return rc;
	size_t f = 0, t = 0;
	HIP_OK(hipMemGetInfo(&f, &t));
	*free_bytes = f;
	*total_bytes = t;
	return NBL_OK;
}

int nbl_device_copy_probe(nbl_device* dev, uint64_t bytes, float* ms)
{
	clear_error();
	if (!dev || bytes < 16 || !ms)
		return fail(NBL_ERR_ARG, "invalid copy probe arguments");
	int rc = set_device(dev);
	if (rc) return rc;
	const size_t n = (size_t)(bytes / 16);
	float4 *a = nullptr, *b = nullptr;
	if ((rc = dev_alloc(&a, n)) || (rc = dev_alloc(&b, n))) {
		hipFree(a);
		return rc;
	}
	hipEvent_t e0, e1;
	HIP_OK(hipEventCreate(&e0));
	HIP_OK(hipEventCreate(&e1));
	HIP_OK(hipMemsetAsync(a, 0, n * 16, dev->stream));
	HIP_OK(hipEventRecord(e0, dev->stream));
	hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, dev->stream, a, b, n);
	HIP_OK(hipEventRecord(e1, dev->stream));
	HIP_OK(hipEventSynchronize(e1));
	HIP_OK(hipEventElapsedTime(ms, e0, e1));
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	hipFree(a);
	hipFree(b);
	return NBL_OK;
}

// Per-launch timing of k_step with HIP events on the driver's stream.
int nbl_driver_set_timing(nbl_driver* d, int enable)
{
	clear_error();
	if (!d)
		return fail(NBL_ERR_ARG, "null driver");
	d->timing = enable != 0;
	d->step_ms = 0;
	d->step_launches = 0;
	return NBL_OK;
}

int nbl_driver_timing(nbl_driver* d, double* step_ms_total, uint64_t* step_launches)
{
	clear_error();
	if (!d || !step_ms_total || !step_launches)
		return fail(NBL_ERR_ARG, "null argument");
	int rc = read_ctrl(d);
	if (rc) return rc;
	*step_ms_total = d->step_ms;
	*step_launches = d->step_launches;
	return NBL_OK;
}

// Track n primaries (device array) and their cascades to the end with the
// buffers in the fixed device array: no allocation of any kind.
int nbl_track_noalloc(nbl_device* dev, const nbl_geometry* geom, const nbl_material* const* mats,
	int32_t n_mats, uint32_t physics_flags, float energy_threshold, uint64_t seed, uint32_t capacity,
	uint32_t steps_per_launch, const nbl_particle28* primaries, const uint32_t* tags, uint32_t n,
	uint64_t sid_base, nbl_detected* out, uint32_t out_cap, uint32_t* n_detected, uint64_t* steps)
{
	clear_error();
	if (!dev || !geom || !n_detected || (n && !primaries) || (out_cap && !out) || n_mats < 0 ||
		(n_mats > 0 && !mats))
		return fail(NBL_ERR_ARG, "invalid nbl_track_noalloc arguments");
	*n_detected = 0;
	if (steps)
		*steps = 0;
	if (capacity < 64 || capacity > NA_MAX_CAPACITY)
		return fail(NBL_ERR_CAPACITY, "nbl_track_noalloc: capacity must be in [64, " +
			std::to_string(NA_MAX_CAPACITY) + "] (its buffers are a fixed device array)");
	if (n_mats > NA_MAX_MATERIALS)
		return fail(NBL_ERR_CAPACITY, "nbl_track_noalloc: at most " + std::to_string(NA_MAX_MATERIALS) +
			" materials");
	if (steps_per_launch < 1 || steps_per_launch > 1024)
		return fail(NBL_ERR_ARG, "steps_per_launch must be in [1, 1024]");
	if (geom->dev->ordinal != dev->ordinal)
		return fail(NBL_ERR_ARG, "geometry on another device");
	for (int i = 0; i < n_mats; ++i)
		if (!mats[i] || mats[i]->dev->ordinal != dev->ordinal)
			return fail(NBL_ERR_ARG, "material " + std::to_string(i) + " missing or on another device");
	if (geom->max_material >= n_mats)
		return fail(NBL_ERR_ARG, "not enough materials provided for this geometry: expected " +
			std::to_string(geom->max_material + 1) + ", got " + std::to_string(n_mats));
	int rc = set_device(dev);
	if (rc) return rc;
	if (n == 0)
		return NBL_OK;
	char* base = dev->scratch;
	Ctrl* ctrl = (Ctrl*)base;
	Material* dmats = (Material*)(base + NA_OFF_MATS);
	const hipStream_t st = dev->stream;
	for (int i = 0; i < n_mats; ++i)
		HIP_OK(hipMemcpyAsync(dmats + i, &mats[i]->m, sizeof(Material), hipMemcpyHostToDevice, st));
	const uint32_t slots = 2u * capacity;
	StepArgs a = {};
	a.W.g = geom->g;
	a.W.mats = dmats;
	a.W.nmats = n_mats;
	a.W.flags = physics_flags;
	a.W.threshold = energy_threshold;
	a.W.key0 = (uint32_t)seed;
	a.W.key1 = (uint32_t)(seed >> 32);
	a.worlds = nullptr;
	a.buf[0] = { base + NA_OFF_BUF0, nullptr, slots };
	a.buf[1] = { base + NA_OFF_BUF1, nullptr, slots };
	a.inel = { base + NA_OFF_INEL, nullptr, slots };
	a.spill = { base + NA_OFF_SPILL, nullptr, NA_SPILL };
	a.spill_cap = NA_SPILL;
	a.ctrl = ctrl;
	a.det = out;
	a.det_sc = nullptr;
	a.dep = nullptr;
	a.dep_cap = 0;
	a.map.np = 0;
	std::memset(a.map.pidx, -1, sizeof(a.map.pidx));
	a.det_cap = out_cap;
	a.slots = slots;
	a.K = steps_per_launch;
	const uint32_t g_refill = blocks_for(capacity);
	const uint32_t g_step = std::min<uint32_t>(blocks_for(capacity), 64u * (uint32_t)dev->n_cu);
	(void)hipGetLastError();   // not this call's
	hipLaunchKernelGGL(k_na_init, dim3(1), dim3(64), 0, st, ctrl, capacity, (uint64_t)n);
	Ctrl h;
	uint32_t frame = 4;
	uint64_t last_steps = ~0ull;
	uint32_t last_det = 0;
	int stalled = 0;
	for (;;) {
		for (uint32_t it = 0; it < frame; ++it) {
			hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(64), 0, st, ctrl, slots, NA_SPILL);
			hipLaunchKernelGGL(k_refill, dim3(g_refill), dim3(256), 0, st, ctrl, a.buf[0], a.buf[1], a.spill,
				primaries, tags, (const uint64_t*)nullptr, (const uint16_t*)nullptr, 0u, sid_base);
			hipLaunchKernelGGL((k_step<false, false>), dim3(g_step), dim3(256), 0, st, a);
			hipLaunchKernelGGL((k_inel<false, false>), dim3(g_step), dim3(256), 0, st, a);
		}
		HIP_OK(hipGetLastError());
		HIP_OK(hipMemcpyAsync(&h, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
		HIP_OK(hipStreamSynchronize(st));
		if (steps)
			*steps = total_steps(h);
		if (h.det_ctr > out_cap) {
			*n_detected = out_cap;
			return fail(NBL_ERR_CAPACITY, "nbl_track_noalloc: more electrons detected than the output "
				"array holds (" + std::to_string(out_cap) + ")");
		}
		*n_detected = h.det_ctr;
		if (active_count(h, NA_SPILL) == 0 && h.queue_n == h.queue_next)
			break;
		// the spill stack cannot grow here: a run whose buffer and stack are
		// full of particles waiting for secondary slots ends with an error
		if (total_steps(h) == last_steps && h.det_ctr == last_det) {
			if (++stalled >= 4)
				return fail(NBL_ERR_CAPACITY, "nbl_track_noalloc: no progress: particle buffer and spill "
					"stack full of particles waiting for secondary-electron slots");
		} else {
			stalled = 0;
		}
		last_steps = total_steps(h);
		last_det = h.det_ctr;
		frame = std::min<uint32_t>(frame * 2, 32);
	}
	return NBL_OK;
}

} // extern "C"
