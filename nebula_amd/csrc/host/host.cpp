// Host-side half of libnebula.so: error state, file formats, the octree
// builder and material preprocessing.  No GPU code here; the HIP half is in
// ../hip/nebula.hip.  Reference paths are relative to its source/ directory.
#include "host.hpp"

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <queue>
#include <unordered_map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

namespace nbl {

// ------------------------------------------------------------------ errors --
static thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
	g_last_error = msg;
	return code;
}

void clear_error() { g_last_error.clear(); }

// ------------------------------------------------------------- vector math --
// Expression orders of common/vector_math.h.
static inline V3 vadd(V3 a, V3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
static inline V3 vsub(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
static inline V3 smul(float a, V3 b) { return { a * b.x, a * b.y, a * b.z }; }
static inline bool veq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

static inline V3 tri_r0(const nbl_triangle44& t) { return { t.r0[0], t.r0[1], t.r0[2] }; }
static inline V3 tri_e1(const nbl_triangle44& t) { return { t.e1[0], t.e1[1], t.e1[2] }; }
static inline V3 tri_e2(const nbl_triangle44& t) { return { t.e2[0], t.e2[1], t.e2[2] }; }
static inline V3 tri_r1(const nbl_triangle44& t) { return vadd(tri_r0(t), tri_e1(t)); }
static inline V3 tri_r2(const nbl_triangle44& t) { return vadd(tri_r0(t), tri_e2(t)); }

// ---------------------------------------------------------- octree builder --
namespace {

// Separating-axis triangle/box test of T. Akenine-Moller ("Fast 3D
// triangle-box overlap testing", JGT 2001) in double precision, with the
// vertex pair and product order of geometry/octree/tribox.cc:109-184 for each
// of the nine edge axes.
struct TriBox {
	double v[3][3], e[3][3], h[3];

	bool edge_axis(int k, int axis) const
	{
		// axis 0: (0, e.z, -e.y); 1: (-e.z, 0, e.x); 2: (e.y, -e.x, 0)
		static const int pair[3][3][2] = {
			{ { 0, 2 }, { 0, 2 }, { 1, 2 } },   // edge 0
			{ { 0, 2 }, { 0, 2 }, { 0, 1 } },   // edge 1
			{ { 0, 1 }, { 0, 1 }, { 1, 2 } } }; // edge 2
		const double* E = e[k];
		const double* A = v[pair[k][axis][0]];
		const double* B = v[pair[k][axis][1]];
		double pa, pb, rad;
		if (axis == 0) {
			pa = E[2] * A[1] - E[1] * A[2];
			pb = E[2] * B[1] - E[1] * B[2];
			rad = std::fabs(E[2]) * h[1] + std::fabs(E[1]) * h[2];
		} else if (axis == 1) {
			pa = -E[2] * A[0] + E[0] * A[2];
			pb = -E[2] * B[0] + E[0] * B[2];
			rad = std::fabs(E[2]) * h[0] + std::fabs(E[0]) * h[2];
		} else {
			pa = E[1] * A[0] - E[0] * A[1];
			pb = E[1] * B[0] - E[0] * B[1];
			rad = std::fabs(E[1]) * h[0] + std::fabs(E[0]) * h[1];
		}
		const double lo = std::min(pa, pb), hi = std::max(pa, pb);
		return !(lo > rad || hi < -rad);
	}

	bool overlap(const double c[3], const double half[3], const double tv[3][3])
	{
		for (int q = 0; q < 3; ++q) {
			h[q] = half[q];
			for (int i = 0; i < 3; ++i)
				v[i][q] = tv[i][q] - c[q];
		}
		for (int q = 0; q < 3; ++q) {
			e[0][q] = v[1][q] - v[0][q];
			e[1][q] = v[2][q] - v[1][q];
			e[2][q] = v[0][q] - v[2][q];
		}
		for (int q = 0; q < 3; ++q) {
			double lo = v[0][q], hi = v[0][q];
			for (int i = 1; i < 3; ++i) {
				if (v[i][q] < lo) lo = v[i][q];
				if (v[i][q] > hi) hi = v[i][q];
			}
			if (lo > h[q] || hi < -h[q])
				return false;
		}
		double n[3] = {
			e[0][1] * e[1][2] - e[0][2] * e[1][1],
			e[0][2] * e[1][0] - e[0][0] * e[1][2],
			e[0][0] * e[1][1] - e[0][1] * e[1][0] };
		double vmin[3], vmax[3];
		for (int q = 0; q < 3; ++q) {
			if (n[q] > 0) { vmin[q] = -h[q] - v[0][q]; vmax[q] = h[q] - v[0][q]; }
			else { vmin[q] = h[q] - v[0][q]; vmax[q] = -h[q] - v[0][q]; }
		}
		if (n[0] * vmin[0] + n[1] * vmin[1] + n[2] * vmin[2] > 0)
			return false;
		if (!(n[0] * vmax[0] + n[1] * vmax[1] + n[2] * vmax[2] >= 0))
			return false;
		for (int k = 0; k < 3; ++k)
			for (int axis = 0; axis < 3; ++axis)
				if (!edge_axis(k, axis))
					return false;
		return true;
	}
};

// octree_builder.cpp:30-39: the cell is inflated by 1.01 in double.
bool cell_overlap(const nbl_triangle44& t, V3 c, V3 h)
{
	const double center[3] = { c.x, c.y, c.z };
	const double half[3] = { 1.01 * (double)h.x, 1.01 * (double)h.y, 1.01 * (double)h.z };
	const V3 a = tri_r0(t), b = tri_r1(t), d = tri_r2(t);
	const double tv[3][3] = { { a.x, a.y, a.z }, { b.x, b.y, b.z }, { d.x, d.y, d.z } };
	TriBox tb;
	return tb.overlap(center, half, tv);
}


struct Node {
	bool leaf = true;
	int child[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
	std::vector<int32_t> tris;
};

struct Builder {
	const nbl_triangle44* tris;
	std::vector<Node> nodes;
	static constexpr int max_depth = 21;   // octree_builder.h:241
	static constexpr size_t split_count = 16;

	// octree_builder.cpp:48-65
	static void octant(int oct, V3 c, V3 h, V3& oc, V3& oh)
	{
		oh = smul(0.5f, h);
		oc.x = c.x + (float)(((oct & 1) << 1) - 1) * oh.x;
		oc.y = c.y + (float)(((oct & 2)) - 1) * oh.y;
		oc.z = c.z + (float)(((oct & 4) >> 1) - 1) * oh.z;
	}

	// A node at depth d holding the duplicate-free, index-ordered triangles
	// overlapping it and its ancestors splits iff it holds more than 16 and
	// d < 21 -- the fixed point of the one-at-a-time insertion of
	// octree_builder.cpp:93-155 (a node's set only grows, leaf lists keep
	// insertion order).  Duplicates are removed once, before the build
	// (see dedupe).
	void build(int id, V3 c, V3 h, int depth, const std::vector<int32_t>& cand)
	{
		std::vector<int32_t> mine;
		mine.reserve(cand.size());
		for (int32_t ti : cand)
			if (cell_overlap(tris[ti], c, h))
				mine.push_back(ti);
		if (mine.size() > split_count && depth < max_depth) {
			nodes[id].leaf = false;
			for (int o = 0; o < 8; ++o) {
				nodes[id].child[o] = (int)nodes.size();
				nodes.emplace_back();
			}
			for (int o = 0; o < 8; ++o) {
				V3 oc, oh;
				octant(o, c, h, oc, oh);
				build(nodes[id].child[o], oc, oh, depth + 1, mine);
			}
		} else {
			nodes[id].tris = std::move(mine);
		}
	}
};

// octree_builder.cpp:100-111 drops a triangle inserted into a leaf that
// already holds a perfect overlap of it (core/triangle.inl:86-103: the same
// three vertices in any order).  Identical triangles overlap the same cells,
// so this keeps, everywhere, the first triangle (by index) of each group of
// perfect overlaps: one pass with a hash of the sorted vertex triple instead
// of comparisons within every node.
struct VKey {
	float v[9];
	bool operator==(const VKey& o) const
	{
		for (int k = 0; k < 9; ++k)
			if (!(v[k] == o.v[k]))
				return false;
		return true;
	}
};
struct VKeyHash {
	size_t operator()(const VKey& k) const
	{
		uint64_t h = 1469598103934665603ull;
		for (float f : k.v) {
			uint32_t u;
			std::memcpy(&u, &f, 4);
			h = (h ^ u) * 1099511628211ull;
		}
		return (size_t)h;
	}
};

std::vector<uint8_t> dedupe(const nbl_triangle44* tris, int32_t ntri)
{
	std::vector<uint8_t> keep((size_t)ntri, 1);
	std::unordered_map<VKey, int32_t, VKeyHash> first;
	first.reserve((size_t)ntri * 2);
	for (int32_t i = 0; i < ntri; ++i) {
		V3 p[3] = { tri_r0(tris[i]), tri_r1(tris[i]), tri_r2(tris[i]) };
		bool nan = false;
		for (V3& q : p) {
			// -0 == +0 under ==; NaN equals nothing, so such a triangle is unique
			if (q.x != q.x || q.y != q.y || q.z != q.z) nan = true;
			if (q.x == 0.0f) q.x = 0.0f;
			if (q.y == 0.0f) q.y = 0.0f;
			if (q.z == 0.0f) q.z = 0.0f;
		}
		if (nan)
			continue;
		std::sort(p, p + 3, [](const V3& a, const V3& b) {
			return a.x < b.x || (a.x == b.x && (a.y < b.y || (a.y == b.y && a.z < b.z)));
		});
		VKey k;
		for (int j = 0; j < 3; ++j) {
			k.v[3 * j] = p[j].x; k.v[3 * j + 1] = p[j].y; k.v[3 * j + 2] = p[j].z;
		}
		if (!first.emplace(k, i).second) {
			keep[(size_t)i] = 0;
			std::fprintf(stderr, "WARNING: ignoring duplicate triangle!\n");
		}
	}
	return keep;
}

} // namespace

int build_octree(const nbl_triangle44* tris, int32_t ntri, std::vector<int32_t>& od,
	std::vector<nbl_triangle44>& sorted, float center[3], float halfsize[3])
{
	if (ntri <= 0)
		return fail(NBL_ERR_ARG, "Octree needs at least one triangle");
	// root box: triangles' AABB, halfsize + 1 nm (octree_builder.cpp:171-209)
	V3 mn{}, mx{};
	for (int32_t i = 0; i < ntri; ++i) {
		const V3 A = tri_r0(tris[i]), B = tri_r1(tris[i]), C = tri_r2(tris[i]);
		const V3 lo{ std::min(std::min(A.x, B.x), C.x), std::min(std::min(A.y, B.y), C.y), std::min(std::min(A.z, B.z), C.z) };
		const V3 hi{ std::max(std::max(A.x, B.x), C.x), std::max(std::max(A.y, B.y), C.y), std::max(std::max(A.z, B.z), C.z) };
		if (i == 0) { mn = lo; mx = hi; continue; }
		mn = { std::min(mn.x, lo.x), std::min(mn.y, lo.y), std::min(mn.z, lo.z) };
		mx = { std::max(mx.x, hi.x), std::max(mx.y, hi.y), std::max(mx.z, hi.z) };
	}
	const V3 c = smul(0.5f, vadd(mn, mx));
	V3 h = smul(0.5f, vsub(mx, mn));
	h = vadd(h, V3{ 1.0f, 1.0f, 1.0f });

	Builder b;
	b.tris = tris;
	const std::vector<uint8_t> keep = dedupe(tris, ntri);
	std::vector<int32_t> all;
	for (int32_t i = 0; i < ntri; ++i)
		if (keep[(size_t)i] && cell_overlap(tris[i], c, h))
			all.push_back(i);
	b.nodes.emplace_back();
	b.nodes[0].leaf = false;   // the root always has eight children
	for (int o = 0; o < 8; ++o) {
		b.nodes[0].child[o] = (int)b.nodes.size();
		b.nodes.emplace_back();
	}
	for (int o = 0; o < 8; ++o) {
		V3 oc, oh;
		Builder::octant(o, c, h, oc, oh);
		b.build(b.nodes[0].child[o], oc, oh, 1, all);
	}

	// breadth-first (= Morton) order, triangles renumbered by first leaf
	// appearance, index table with negated leaf offsets
	// (octree_builder.cpp:252-347)
	std::vector<int> order;
	order.reserve(b.nodes.size());
	order.push_back(0);
	for (size_t k = 0; k < order.size(); ++k)
		if (!b.nodes[order[k]].leaf)
			for (int o = 0; o < 8; ++o)
				order.push_back(b.nodes[order[k]].child[o]);
	std::vector<int32_t> map(ntri, -1);
	sorted.clear();
	size_t total = 0;
	for (int id : order) {
		const Node& nd = b.nodes[id];
		if (!nd.leaf) { total += 8; continue; }
		total += nd.tris.size() + 1;
		for (int32_t ti : nd.tris)
			if (map[ti] < 0) {
				map[ti] = (int32_t)sorted.size();
				sorted.push_back(tris[ti]);
			}
	}
	if (total > (size_t)0x7fffffff)
		return fail(NBL_ERR_CAPACITY, "Octree is too large.");
	od.assign(total, 0);
	std::vector<int32_t> start(b.nodes.size());
	size_t pos = 0;
	for (int id : order) {
		const Node& nd = b.nodes[id];
		start[id] = (int32_t)pos;
		if (nd.leaf) {
			for (int32_t ti : nd.tris)
				od[pos++] = map[ti];
			od[pos++] = -1;
		} else {
			pos += 8;
		}
	}
	for (int id : order) {
		const Node& nd = b.nodes[id];
		if (nd.leaf) continue;
		for (int o = 0; o < 8; ++o) {
			const int cid = nd.child[o];
			od[start[id] + o] = b.nodes[cid].leaf ? -start[cid] : start[cid];
		}
	}
	center[0] = c.x; center[1] = c.y; center[2] = c.z;
	halfsize[0] = h.x; halfsize[1] = h.y; halfsize[2] = h.z;
	return NBL_OK;
}

// ------------------------------------------------------------ materials --
namespace {

// common/util/range.h: geomspace[i] = first * exp(i * log(stop/start)/(n-1)),
// linspace[i] = first + i * (stop - start)/(n-1).
double geom_at(double front, double back, int n, int i)
{
	const double step = std::log(back / front) / (double)(size_t)(n - 1);
	return front * std::exp((double)(size_t)i * step);
}
double lin_at(double front, double back, int n, int i)
{
	const double step = (back - front) / (double)(size_t)(n - 1);
	return front + (double)(size_t)i * step;
}

bool alloc_table(nbl_table& t, const nbl_raw_dataset& d, int ndim)
{
	if (!d.data || d.ndim != ndim)
		return false;
	t.ndim = ndim;
	size_t count = 1;
	for (int a = 0; a < ndim; ++a) {
		t.dims[a] = d.dims[a];
		if (d.dims[a] < 2)
			return false;
		count *= (size_t)d.dims[a];
	}
	t.data = (float*)std::malloc(sizeof(float) * count);
	if (!t.data)
		return false;
	std::memcpy(t.data, d.data, sizeof(float) * count);
	return true;
}

size_t table_count(const nbl_table& t)
{
	size_t c = 1;
	for (int a = 0; a < t.ndim; ++a)
		c *= (size_t)t.dims[a];
	return c;
}

// table set_scale: step = (n-1)/(max-min) in float.
void log_axis(nbl_table& t, const nbl_raw_dataset& d, int a)
{
	const int n = t.dims[a];
	const float lo = (float)std::log(d.ax_front[a]);
	const float hi = (float)std::log(geom_at(d.ax_front[a], d.ax_back[a], n, n - 1));
	t.min[a] = lo;
	t.step[a] = (float)(n - 1) / (hi - lo);
}
void lin_range(const nbl_raw_dataset& d, int a, int n, double& front, double& back_raw)
{
	if (d.ax_present[a]) { front = d.ax_front[a]; back_raw = d.ax_back[a]; }
	else { front = 0.0; back_raw = 1.0; }
	(void)n;
}
void lin_axis(nbl_table& t, const nbl_raw_dataset& d, int a)
{
	const int n = t.dims[a];
	double f, b;
	lin_range(d, a, n, f, b);
	const float lo = (float)f, hi = (float)lin_at(f, b, n, n - 1);
	t.min[a] = lo;
	t.step[a] = (float)(n - 1) / (hi - lo);
}

// log(imfp) (physics/kieft/elastic.h:128-141 and the same block of
// full_penn.h:305-318 / kieft/inelastic.h:222-235).
bool make_log_imfp(nbl_table& t, const nbl_raw_dataset& d)
{
	if (!alloc_table(t, d, 1)) return false;
	const float unit = (float)(d.unit * 1.0);
	for (int x = 0; x < t.dims[0]; ++x)
		t.data[x] = std::log(t.data[x] * unit);
	log_axis(t, d, 0);
	return true;
}

// log(omega) clamped to K - fermi (full_penn.h:320-345,
// kieft/inelastic.h:237-264).
bool make_log_omega(nbl_table& t, const nbl_raw_dataset& d, double fermi)
{
	if (!alloc_table(t, d, 2)) return false;
	const int nx = t.dims[0], ny = t.dims[1];
	const float unit = (float)(d.unit / 1.0);
	for (int x = 0; x < nx; ++x) {
		const double K = geom_at(d.ax_front[0], d.ax_back[0], nx, x);
		for (int y = 0; y < ny; ++y) {
			float& v = t.data[(size_t)x * ny + y];
			v = (float)std::log(std::max(0.0, std::min<double>((K - fermi) / 1.0, v * unit)));
		}
	}
	log_axis(t, d, 0);
	lin_axis(t, d, 1);
	return true;
}

float get_rounddown2(const nbl_table& t, float x, float y)
{
	const float xi = (x - t.min[0]) * t.step[0];
	const float yi = (y - t.min[1]) * t.step[1];
	const int lx = std::min(std::max((int)std::floor(xi), 0), t.dims[0] - 1);
	const int ly = std::min(std::max((int)std::floor(yi), 0), t.dims[1] - 1);
	return t.data[(size_t)lx * t.dims[1] + ly];
}

void free_material(nbl_material_host* m)
{
	if (!m) return;
	nbl_table* ts[] = { &m->el_imfp, &m->el_icdf, &m->in_imfp, &m->in_icdf, &m->pn_q, &m->pn_ion, &m->ki_ion };
	for (nbl_table* t : ts)
		std::free(t->data);
	std::free(m);
}

} // namespace

int prepare_material(const nbl_raw_material* r, nbl_material_host** out)
{
	if (!r || !out)
		return fail(NBL_ERR_ARG, "null argument");
	auto* m = (nbl_material_host*)std::calloc(1, sizeof(nbl_material_host));
	if (!m)
		return fail(NBL_ERR_ALLOC, "out of host memory");
	auto bad = [&](const char* what) {
		free_material(m);
		return fail(NBL_ERR_FORMAT, std::string("Material table missing or malformed: ") + what);
	};
	// core/material.inl:13-18
	m->barrier = (float)(r->barrier / 1.0);
	m->inel_model = r->inel_model;
	// kieft_elastic::create (physics/kieft/elastic.h:120-166)
	if (!make_log_imfp(m->el_imfp, r->el_imfp)) return bad("/elastic/imfp");
	if (!alloc_table(m->el_icdf, r->el_icdf, 2)) return bad("/elastic/costheta_icdf");
	for (size_t i = 0; i < table_count(m->el_icdf); ++i)
		m->el_icdf.data[i] = std::max(-1.0f, std::min(1.0f, m->el_icdf.data[i]));
	log_axis(m->el_icdf, r->el_icdf, 0);
	lin_axis(m->el_icdf, r->el_icdf, 1);
	m->phonon_loss = (float)(r->phonon_loss / 1.0);
	m->recoil_const = 2 * (float)((2 * (9.109383e-28 * 6.2415e15)) / r->effective_A);
	m->fermi = (float)(r->fermi / 1.0);
	m->band_gap = (float)((r->has_band_gap ? r->band_gap : -1.0) / 1.0);

	if (r->inel_model == NBL_INEL_PENN) {
		// full_penn::create (physics/full_penn.h:294-406)
		if (!make_log_imfp(m->in_imfp, r->in_imfp)) return bad("/full_penn/imfp");
		if (!make_log_omega(m->in_icdf, r->in_icdf, r->fermi)) return bad("/full_penn/omega_icdf");
		if (!alloc_table(m->pn_q, r->pn_q, 3)) return bad("/full_penn/q_icdf");
		{
			const float unit = (float)(r->pn_q.unit * 0.19519 * 1.0);
			for (size_t i = 0; i < table_count(m->pn_q); ++i)
				m->pn_q.data[i] = m->pn_q.data[i] * unit;
			log_axis(m->pn_q, r->pn_q, 0);
			lin_axis(m->pn_q, r->pn_q, 1);
			lin_axis(m->pn_q, r->pn_q, 2);
		}
		if (!alloc_table(m->pn_ion, r->pn_ion, 3)) return bad("/ionization/binding_icdf (3-D)");
		{
			const float unit = (float)(r->pn_ion.unit / 1.0);
			for (size_t i = 0; i < table_count(m->pn_ion); ++i) {
				float b = m->pn_ion.data[i] * unit;
				if (b < 50 || !std::isfinite(b))
					b = -1;
				m->pn_ion.data[i] = b;
			}
			log_axis(m->pn_ion, r->pn_ion, 0);
			log_axis(m->pn_ion, r->pn_ion, 1);
			lin_axis(m->pn_ion, r->pn_ion, 2);
		}
	} else if (r->inel_model == NBL_INEL_KIEFT) {
		// kieft_inelastic::create (physics/kieft/inelastic.h:212-327)
		if (!make_log_imfp(m->in_imfp, r->in_imfp)) return bad("/inelastic_kieft/imfp");
		if (!make_log_omega(m->in_icdf, r->in_icdf, r->fermi)) return bad("/inelastic_kieft/w0_icdf");
		const nbl_raw_dataset& d = r->ki_ion;
		if (!alloc_table(m->ki_ion, d, 2)) return bad("/ionization/binding_icdf (2-D)");
		nbl_table& t = m->ki_ion;
		log_axis(t, d, 0);
		lin_axis(t, d, 1);
		std::vector<double> shells;
		for (int i = 0; i < r->n_outer_shells; ++i) {
			const double v = r->outer_shells[i] * r->outer_shells_unit;
			if (v < 100.0)
				shells.push_back(v);
		}
		std::sort(shells.begin(), shells.end());
		const int nx = t.dims[0], ny = t.dims[1];
		double Pf, Pb;
		lin_range(d, 1, ny, Pf, Pb);
		// in-place rebuild reading its own partly rewritten rows (:287-321)
		for (int x = 0; x < nx; ++x) {
			const double K = geom_at(d.ax_front[0], d.ax_back[0], nx, x);
			for (int y = 0; y < ny; ++y) {
				const float P = (float)lin_at(Pf, Pb, ny, y);
				double binding = -1;
				if (K > 100) {
					binding = (double)get_rounddown2(t, (float)std::log((K + 10.0) / 1.0), P) * d.unit;
					if (binding < 50 || !std::isfinite(binding))
						binding = -1;
				}
				if (binding < 0) {
					auto it = std::lower_bound(shells.rbegin(), shells.rend(), K, std::greater<double>{});
					if (it != shells.rend())
						binding = *it;
				}
				t.data[(size_t)x * ny + y] = (float)(binding / 1.0);
			}
		}
	} else {
		free_material(m);
		return fail(NBL_ERR_ARG, "unknown inelastic model");
	}
	*out = m;
	return NBL_OK;
}

float material_max_energy(const nbl_material_host* m)
{
	// exp(get_scalemax()) of both IMFP tables, minimum over mechanisms
	auto smax = [](const nbl_table& t) { return t.min[0] + (float)(t.dims[0] - 1) / t.step[0]; };
	float e = INFINITY;
	const float el = std::exp(smax(m->el_imfp));
	const float in = std::exp(smax(m->in_imfp));
	if (el < e) e = el;
	if (in < e) e = in;
	return e;
}

// ---------------------------------------------------------------- formats --
int load_tri_file(const char* path, std::vector<nbl_triangle44>& out)
{
	std::ifstream ifs(path);
	if (!ifs.is_open())
		return fail(NBL_ERR_IO, std::string("Unable to open file ") + path);
	out.clear();
	size_t line_num = 1;
	std::string line;
	while (std::getline(ifs, line)) {
		++line_num;
		if (line.empty())
			continue;
		const size_t i = line.find_first_not_of(" \t");
		if (i < line.size() && line[i] == '#')
			continue;
		std::istringstream ss(line);
		std::vector<std::string> cols;
		std::string c;
		while (ss >> c)
			cols.push_back(c);
		if (cols.empty())
			continue;
		if (cols.size() != 11) {
			std::ostringstream oss;
			oss << "invalid number of columns in line '" << line_num << "'";
			return fail(NBL_ERR_FORMAT, oss.str());
		}
		try {
			float v[9];
			for (int k = 0; k < 9; ++k)
				v[k] = (float)std::stod(cols[2 + k]);
			nbl_triangle44 t;
			t.material_in = std::stoi(cols[0]);
			t.material_out = std::stoi(cols[1]);
			for (int k = 0; k < 3; ++k) {
				t.r0[k] = v[k];
				t.e1[k] = v[3 + k] - v[k];
				t.e2[k] = v[6 + k] - v[k];
			}
			out.push_back(t);
		} catch (const std::exception& e) {
			std::ostringstream oss;
			oss << "cannot parse line '" << line_num << "': " << e.what();
			return fail(NBL_ERR_FORMAT, oss.str());
		}
	}
	return NBL_OK;
}

int load_pri_file(const char* path, const float mn[3], const float mx[3], float max_E,
	std::vector<nbl_record36>& out, int64_t warnings[4])
{
	std::ifstream ifs(path, std::ios::binary);
	if (!ifs.is_open())
		return fail(NBL_ERR_IO, std::string("Unable to open file ") + path);
	const float eps = 10 * std::numeric_limits<float>::epsilon();
	for (int k = 0; k < 4; ++k)
		warnings[k] = 0;
	out.clear();
	std::vector<char> buf(36 * 1024);
	for (;;) {
		ifs.read(buf.data(), (std::streamsize)buf.size());
		const size_t got = (size_t)ifs.gcount();
		if (got == 0)
			break;
		if (got % 36)
			return fail(NBL_ERR_FORMAT, "Unexpected end of primaries file");
		for (size_t i = 0; i < got / 36; ++i) {
			nbl_record36 r;
			std::memcpy(&r, buf.data() + 36 * i, 36);
			const float* p = r.pos;
			if (!(p[0] > mn[0] && p[1] > mn[1] && p[2] > mn[2] && p[0] < mx[0] && p[1] < mx[1] && p[2] < mx[2]))
				++warnings[0];
			if (!(r.kin_energy > eps))
				++warnings[1];
			if (r.kin_energy > max_E)
				++warnings[2];
			if (!(std::fabs(r.dir[0]) > eps || std::fabs(r.dir[1]) > eps || std::fabs(r.dir[2]) > eps))
				++warnings[3];
			out.push_back(r);
		}
		if (!ifs)
			break;
	}
	return NBL_OK;
}

void prescan_shuffle(nbl_record36* rec, int64_t n, int64_t prescan_size)
{
	std::default_random_engine rng;
	const size_t N = (size_t)n;
	for (size_t i1 = 0; i1 < std::min((size_t)prescan_size, N); ++i1) {
		std::uniform_int_distribution<size_t> dist(i1, N - 1);
		const size_t i2 = dist(rng);
		std::swap(rec[i1], rec[i2]);
	}
}

// sort_pri_file (io/load_pri_file.cpp:29-40,186-229): order primaries by the
// Morton code of their position quantised to 16 bits per axis over the
// primaries' bounding box.  The reference sorts an index permutation with
// std::sort (not stable) and a comparator that evaluates both codes; sorting
// the same permutation with the same comparator outcomes gives the same order,
// so the codes are computed once.  A float->uint16 conversion is made through
// int32, as x86-64 code does (an axis of zero extent gives 0/0: NaN -> 0).
static uint64_t morton16(uint16_t x, uint16_t y, uint16_t z)
{
	uint64_t m = 0;
	for (unsigned i = 0; i < 16; ++i) {
		m |= (x & (uint64_t(1) << i)) << (2 * i + 0);
		m |= (y & (uint64_t(1) << i)) << (2 * i + 1);
		m |= (z & (uint64_t(1) << i)) << (2 * i + 2);
	}
	return m;
}

static uint16_t quantise16(float v)
{
	return (uint16_t)(int32_t)v;
}

void sort_primaries(nbl_record36* rec, int64_t n)
{
	if (n <= 0)
		return;
	float vmin[3], vmax[3];
	for (int k = 0; k < 3; ++k)
		vmin[k] = vmax[k] = rec[0].pos[k];
	for (int64_t i = 0; i < n; ++i)
		for (int k = 0; k < 3; ++k) {
			vmin[k] = std::min(vmin[k], rec[i].pos[k]);
			vmax[k] = std::max(vmax[k], rec[i].pos[k]);
		}
	float size[3];
	for (int k = 0; k < 3; ++k)
		size[k] = vmax[k] - vmin[k];
	const float qmax = (float)std::numeric_limits<uint16_t>::max();
	std::vector<uint64_t> code((size_t)n);
	for (int64_t i = 0; i < n; ++i) {
		uint16_t q[3];
		for (int k = 0; k < 3; ++k)
			q[k] = quantise16((rec[i].pos[k] - vmin[k]) / size[k] * qmax);
		code[(size_t)i] = morton16(q[0], q[1], q[2]);
	}
	std::vector<size_t> perm((size_t)n);
	for (size_t i = 0; i < perm.size(); ++i)
		perm[i] = i;
	std::sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return code[a] < code[b]; });
	std::vector<nbl_record36> sorted((size_t)n);
	for (size_t i = 0; i < perm.size(); ++i)
		sorted[i] = rec[perm[i]];
	std::memcpy(rec, sorted.data(), sizeof(nbl_record36) * (size_t)n);
}

} // namespace nbl

// =================================================================== C ABI ==
using namespace nbl;

extern "C" {

const char* nbl_last_error(void) { return g_last_error.c_str(); }
int nbl_version(void) { return 100; }
void nbl_free(void* p) { std::free(p); }

int nbl_load_tri_file(const char* path, nbl_triangle44** tris, int32_t* n)
{
	clear_error();
	if (!path || !tris || !n)
		return fail(NBL_ERR_ARG, "null argument");
	std::vector<nbl_triangle44> v;
	const int rc = load_tri_file(path, v);
	if (rc != NBL_OK)
		return rc;
	*tris = (nbl_triangle44*)std::malloc(sizeof(nbl_triangle44) * (v.size() ? v.size() : 1));
	if (!*tris)
		return fail(NBL_ERR_ALLOC, "out of host memory");
	std::memcpy(*tris, v.data(), sizeof(nbl_triangle44) * v.size());
	*n = (int32_t)v.size();
	return NBL_OK;
}

int nbl_load_pri_file(const char* path, const float aabb_min[3], const float aabb_max[3],
	float max_E, nbl_record36** records, int64_t* n, int64_t warnings[4])
{
	clear_error();
	if (!path || !records || !n || !warnings || !aabb_min || !aabb_max)
		return fail(NBL_ERR_ARG, "null argument");
	std::vector<nbl_record36> v;
	const int rc = load_pri_file(path, aabb_min, aabb_max, max_E, v, warnings);
	if (rc != NBL_OK)
		return rc;
	*records = (nbl_record36*)std::malloc(sizeof(nbl_record36) * (v.size() ? v.size() : 1));
	if (!*records)
		return fail(NBL_ERR_ALLOC, "out of host memory");
	std::memcpy(*records, v.data(), sizeof(nbl_record36) * v.size());
	*n = (int64_t)v.size();
	return NBL_OK;
}

int nbl_prescan_shuffle(nbl_record36* records, int64_t n, int64_t prescan_size)
{
	clear_error();
	if (!records || n < 0 || prescan_size < 0)
		return fail(NBL_ERR_ARG, "invalid argument");
	prescan_shuffle(records, n, prescan_size);
	return NBL_OK;
}

int nbl_sort_primaries(nbl_record36* records, int64_t n)
{
	clear_error();
	if ((!records && n) || n < 0)
		return fail(NBL_ERR_ARG, "invalid argument");
	sort_primaries(records, n);
	return NBL_OK;
}

int nbl_octree_build(const nbl_triangle44* tris, int32_t ntri, int32_t** octree_data,
	int32_t* n_octree, nbl_triangle44** sorted_tris, int32_t* n_sorted,
	float center[3], float halfsize[3])
{
	clear_error();
	if (!tris || !octree_data || !n_octree || !sorted_tris || !n_sorted || !center || !halfsize)
		return fail(NBL_ERR_ARG, "null argument");
	std::vector<int32_t> od;
	std::vector<nbl_triangle44> st;
	const int rc = build_octree(tris, ntri, od, st, center, halfsize);
	if (rc != NBL_OK)
		return rc;
	*octree_data = (int32_t*)std::malloc(sizeof(int32_t) * od.size());
	*sorted_tris = (nbl_triangle44*)std::malloc(sizeof(nbl_triangle44) * (st.size() ? st.size() : 1));
	if (!*octree_data || !*sorted_tris)
		return fail(NBL_ERR_ALLOC, "out of host memory");
	std::memcpy(*octree_data, od.data(), sizeof(int32_t) * od.size());
	std::memcpy(*sorted_tris, st.data(), sizeof(nbl_triangle44) * st.size());
	*n_octree = (int32_t)od.size();
	*n_sorted = (int32_t)st.size();
	return NBL_OK;
}

int nbl_material_prepare(const nbl_raw_material* raw, nbl_material_host** out)
{
	clear_error();
	return prepare_material(raw, out);
}

void nbl_material_host_free(nbl_material_host* m) { free_material(m); }

float nbl_material_max_energy(const nbl_material_host* m) { return m ? material_max_energy(m) : 0.0f; }

int nbl_material_load_file(const char* path, int32_t inel_model, nbl_material_host** out)
{
	clear_error();
	if (!path || !out)
		return fail(NBL_ERR_ARG, "null argument");
	// The HDF5 reader lives in libnbl_hdf5.so next to this library, so that
	// libnebula.so itself does not depend on the HDF5 C library.
	Dl_info info;
	std::string dir = ".";
	if (dladdr((void*)&nbl_material_load_file, &info) && info.dli_fname) {
		dir = info.dli_fname;
		const size_t s = dir.rfind('/');
		dir = (s == std::string::npos) ? "." : dir.substr(0, s);
	}
	void* h = dlopen((dir + "/libnbl_hdf5.so").c_str(), RTLD_NOW | RTLD_LOCAL);
	if (!h)
		return fail(NBL_ERR_UNSUPPORTED, std::string("HDF5 reader unavailable: ") + dlerror());
	using read_fn = int (*)(const char*, int32_t, nbl_raw_material*, void**, char*, size_t);
	using free_fn = void (*)(void*);
	auto rd = (read_fn)dlsym(h, "nbl_h5_read_material");
	auto fr = (free_fn)dlsym(h, "nbl_h5_free_material");
	if (!rd || !fr)
		return fail(NBL_ERR_UNSUPPORTED, "HDF5 reader lacks its entry points");
	nbl_raw_material raw;
	void* storage = nullptr;
	char msg[512] = { 0 };
	int rc = rd(path, inel_model, &raw, &storage, msg, sizeof(msg));
	if (rc != NBL_OK)
		return fail(rc, msg);
	rc = prepare_material(&raw, out);
	fr(storage);
	return rc;
}

} // extern "C"
