// Harness over the reference's CPU code, compiled from the reference sources
// where they lie (never copied).  It writes the vectors that pin the oracle:
//   octree    <tris.bin> <out.bin>            builder output
//   propagate <tris.bin> <rays.bin> <out.bin> octree<false>::propagate
//   intersect <tris.bin> <rays.bin> <out.bin> triangle::intersect_ray
//   material  <file.mat> <penn|kieft> <out.bin>            prepared tables
//   probe     <file.mat> <penn|kieft> <pts.bin> <out.bin>  table lookups
//   step      <tris.bin> <penn|kieft> <seed> <states.bin> <out.bin> <mat>...
//             one cpu_driver init/intersect/scatter per state, each with
//             its own mt19937(seed + i)
//   simulate  <tris.bin> <penn|kieft> <seed> <pri.bin> <out.bin> <mat>...
//             simple_cpu_driver over all primaries, one thread
//   edep      <tris.bin> <penn|kieft> <seed> <pri.bin> <out.bin> <mat>...
//             energydeposit_cpu_driver over all primaries, one thread:
//             deposit records (float pos[3], E, dE, uint32 tag) in callback
//             order, int64 count; then detected records as for simulate
//   stepx     <tris.bin> <variant> <threshold> <seed> <states.bin> <out.bin> <mat>...
//   simulatex <tris.bin> <variant> <threshold> <seed> <pri.bin> <out.bin> <mat>...
//             step / simulate with the physics of a named variant (the
//             template flags of physics_config.h set otherwise, see
//             `variant` below) and an energy threshold (eV)
//   pri       <in.pri> <prescan_size> <sort 0|1> <out.bin>
//             load_pri_file, optional sort_pri_file, prescan_shuffle, in the
//             order of gpu_main.cu:364-367; out = the 36-byte records
// Binary layouts are little-endian and documented beside each reader.
// Access to private members (tables, particle vector) is needed to dump
// them; the reference files themselves are untouched.
#define private public
#define protected public
#include "config/config.h"
#include "physics_config.h"
#include "core/material.h"
#include "drivers/cpu/simple_cpu_driver.h"
#include "drivers/cpu/energydeposit_cpu_driver.h"
#include "geometry/octree.h"
#include "io/load_pri_file.h"
#undef private
#undef protected

#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

using penn_list = scatter_list<nbl::scatter::full_penn<false, true>, nbl::scatter::kieft_elastic<false, true, true>>;
using kieft_list = scatter_list<nbl::scatter::kieft_inelastic<false, true, true, true, true>,
	nbl::scatter::kieft_elastic<false, true, true>>;
using geometry_t = nbl::geometry::octree<false>;

// Non-default physics of the stepx / simulatex commands.  Each names the
// template flags it sets (physics/full_penn.h:18-20, kieft/inelastic.h:
// 18-24, kieft/elastic.h:20-22, boundary_intersect.h:10-14):
//   penn_abs   the default Penn list with interface absorption on
//              (boundary_intersect<true, true, true>);
//   penn_alt   full_penn<no secondaries>, kieft_elastic<phonon loss, no
//              recoil>, boundary_intersect<quantum, no refraction, absorption>;
//   kieft_alt  kieft_inelastic<no optical phonon, secondaries, no
//              instantaneous momentum, no momentum conservation>,
//              kieft_elastic<no phonon loss, no recoil>,
//              boundary_intersect<classical, no refraction, no absorption>.
using intersect_abs = boundary_intersect<true, true, true>;
using penn_alt_list = scatter_list<nbl::scatter::full_penn<false, false>,
	nbl::scatter::kieft_elastic<false, true, false>>;
using intersect_penn_alt = boundary_intersect<true, false, true>;
using kieft_alt_list = scatter_list<nbl::scatter::kieft_inelastic<false, false, true, false, false>,
	nbl::scatter::kieft_elastic<false, false, false>>;
using intersect_kieft_alt = boundary_intersect<false, false, false>;

template <typename T>
static std::vector<T> read_vec(const char* path)
{
	std::ifstream f(path, std::ios::binary | std::ios::ate);
	if (!f)
		throw std::runtime_error(std::string("cannot open ") + path);
	const size_t n = (size_t)f.tellg();
	f.seekg(0);
	std::vector<T> v(n / sizeof(T));
	f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
	return v;
}

struct Out {
	std::ofstream f;
	explicit Out(const char* p) : f(p, std::ios::binary) {}
	template <typename T> void put(const T& v) { f.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
	template <typename T> void put(const T* v, size_t n) { f.write(reinterpret_cast<const char*>(v), sizeof(T) * n); }
};

// tris.bin: per triangle int32 mat_in, mat_out, float32 r0[3], r1[3], r2[3]
// (44 bytes); the triangle is built with the reference constructor.
static std::vector<triangle> read_tris(const char* path)
{
	struct Raw { int32_t in, out; float v[9]; };
	static_assert(sizeof(Raw) == 44, "layout");
	std::vector<Raw> raw = read_vec<Raw>(path);
	std::vector<triangle> t;
	for (const Raw& r : raw)
		t.push_back(triangle({ r.v[0], r.v[1], r.v[2] }, { r.v[3], r.v[4], r.v[5] }, { r.v[6], r.v[7], r.v[8] }, r.in, r.out));
	return t;
}

// rays.bin: float pos[3], dir[3], dist, int32 ignore_tri, ignore_mat, tri
struct Ray { float pos[3], dir[3], dist; int32_t ignore_tri, ignore_mat, tri; };

static int cmd_octree(char** a)
{
	auto tris = read_tris(a[0]);
	nbl::geometry::octree_builder::linearized_octree lin{ nbl::geometry::octree_builder::octree_root(tris) };
	Out o(a[1]);
	o.put((int64_t)lin.octree_data.size());
	o.put(lin.octree_data.data(), lin.octree_data.size());
	o.put((int64_t)lin.triangle_data.size());
	for (const triangle& t : lin.triangle_data)
		o.put(reinterpret_cast<const char*>(&t), sizeof(triangle));
	o.put(lin.center);
	o.put(lin.halfsize);
	return 0;
}

static int cmd_propagate(char** a)
{
	auto tris = read_tris(a[0]);
	auto rays = read_vec<Ray>(a[1]);
	geometry_t g = geometry_t::create(tris);
	Out o(a[2]);
	for (const Ray& r : rays) {
		const triangle* ign = r.ignore_tri >= 0 ? g._triangles + r.ignore_tri : nullptr;
		intersect_event ev = g.propagate({ r.pos[0], r.pos[1], r.pos[2] }, { r.dir[0], r.dir[1], r.dir[2] },
			r.dist, ign, r.ignore_mat);
		const int32_t idx = ev.isect_triangle ? (int32_t)(ev.isect_triangle - g._triangles) : -1;
		o.put(idx);
		o.put(ev.isect_distance);
	}
	o.put(g._max_extent);
	return 0;
}

static int cmd_intersect(char** a)
{
	auto tris = read_tris(a[0]);
	auto rays = read_vec<Ray>(a[1]);
	Out o(a[2]);
	for (const Ray& r : rays) {
		const float t = tris[r.tri].intersect_ray({ r.pos[0], r.pos[1], r.pos[2] }, { r.dir[0], r.dir[1], r.dir[2] });
		o.put(t);
	}
	return 0;
}

template <typename T> static void dump1(Out& o, const nbl::util::table_1D<T, false>& t)
{
	o.put((int32_t)1); o.put((int32_t)t._n); o.put((int32_t)0); o.put((int32_t)0);
	o.put(t._x_min); o.put(t._x_step); o.put(0.f); o.put(0.f); o.put(0.f); o.put(0.f);
	o.put(t._data, (size_t)t._n);
}
template <typename T> static void dump2(Out& o, const nbl::util::table_2D<T, false>& t)
{
	o.put((int32_t)2); o.put((int32_t)t._width); o.put((int32_t)t._height); o.put((int32_t)0);
	o.put(t._x_min); o.put(t._x_step); o.put(t._y_min); o.put(t._y_step); o.put(0.f); o.put(0.f);
	for (int i = 0; i < t._width; ++i)
		for (int j = 0; j < t._height; ++j)
			o.put(t(i, j));
}
template <typename T> static void dump3(Out& o, const nbl::util::table_3D<T, false>& t)
{
	o.put((int32_t)3); o.put((int32_t)t._width); o.put((int32_t)t._height); o.put((int32_t)t._depth);
	o.put(t._x_min); o.put(t._x_step); o.put(t._y_min); o.put(t._y_step); o.put(t._z_min); o.put(t._z_step);
	for (int i = 0; i < t._width; ++i)
		for (int j = 0; j < t._height; ++j)
			for (int k = 0; k < t._depth; ++k)
				o.put(t(i, j, k));
}

// Dump order: barrier, phonon_loss, recoil_const, fermi, band_gap, then the
// tables el_imfp, el_icdf, in_imfp, in_icdf, (penn) pn_q, pn_ion / (kieft) ki_ion.
static void dump_material(Out& o, const material<penn_list>& m)
{
	const auto& el = nbl::tuple::get<1>(m);
	const auto& in = nbl::tuple::get<0>(m);
	o.put(m.barrier); o.put(el._phonon_loss); o.put(el._recoil_const); o.put(in._fermi); o.put(in._band_gap);
	o.put(m.get_max_energy());
	dump1(o, el._log_imfp_table); dump2(o, el._icdf_table);
	dump1(o, in._log_imfp_table); dump2(o, in._log_omega_icdf_table);
	dump3(o, in._q_icdf_table); dump3(o, in._ionisation_table);
}
static void dump_material(Out& o, const material<kieft_list>& m)
{
	const auto& el = nbl::tuple::get<1>(m);
	const auto& in = nbl::tuple::get<0>(m);
	o.put(m.barrier); o.put(el._phonon_loss); o.put(el._recoil_const); o.put(in._fermi); o.put(in._band_gap);
	o.put(m.get_max_energy());
	dump1(o, el._log_imfp_table); dump2(o, el._icdf_table);
	dump1(o, in._log_imfp_table); dump2(o, in._log_icdf_table);
	dump2(o, in._ionisation_table);
}

template <typename L> static int cmd_material(char** a)
{
	material<L> m{ nbl::hdf5_file(a[0]) };
	Out o(a[2]);
	dump_material(o, m);
	return 0;
}

// pts.bin: float x, y, z.  Output per point: el_imfp.get(x), el_icdf.get(x,y),
// el_icdf.get_rounddown(x,y), el_icdf.get_nearest(x,y), in_imfp.get(x),
// in_icdf.get(x,y), then (penn) q.get(x,y,z), q.get_rounddown(x,y,z),
// ion.get_rounddown(x,y,z) / (kieft) ion.get_rounddown(x,y), ion.get(x,y).
struct Pt { float x, y, z; };
template <typename L> static void probe_point(Out& o, const material<L>& m, const Pt& p);
template <> void probe_point(Out& o, const material<penn_list>& m, const Pt& p)
{
	const auto& el = nbl::tuple::get<1>(m);
	const auto& in = nbl::tuple::get<0>(m);
	o.put(el._log_imfp_table.get(p.x)); o.put(el._icdf_table.get(p.x, p.y));
	o.put(el._icdf_table.get_rounddown(p.x, p.y)); o.put(el._icdf_table.get_nearest(p.x, p.y));
	o.put(in._log_imfp_table.get(p.x)); o.put(in._log_omega_icdf_table.get(p.x, p.y));
	o.put(in._q_icdf_table.get(p.x, p.y, p.z)); o.put(in._q_icdf_table.get_rounddown(p.x, p.y, p.z));
	o.put(in._ionisation_table.get_rounddown(p.x, p.y, p.z));
}
template <> void probe_point(Out& o, const material<kieft_list>& m, const Pt& p)
{
	const auto& el = nbl::tuple::get<1>(m);
	const auto& in = nbl::tuple::get<0>(m);
	o.put(el._log_imfp_table.get(p.x)); o.put(el._icdf_table.get(p.x, p.y));
	o.put(el._icdf_table.get_rounddown(p.x, p.y)); o.put(el._icdf_table.get_nearest(p.x, p.y));
	o.put(in._log_imfp_table.get(p.x)); o.put(in._log_icdf_table.get(p.x, p.y));
	o.put(in._ionisation_table.get_rounddown(p.x, p.y)); o.put(in._ionisation_table.get(p.x, p.y));
}

template <typename L> static int cmd_probe(char** a)
{
	material<L> m{ nbl::hdf5_file(a[0]) };
	auto pts = read_vec<Pt>(a[2]);
	Out o(a[3]);
	for (const Pt& p : pts)
		probe_point(o, m, p);
	return 0;
}

// states.bin: float pos[3], dir[3], E, int32 mat, int32 last_tri (36 bytes).
// Output per state: int32 status after the step, int32 next_scatter,
// int32 material, int32 last_tri, float pos[3], dir[3], E, int32 n_se,
// then for a secondary: float pos[3], dir[3], E, int32 material.
struct St { float pos[3], dir[3], E; int32_t mat, last_tri; };

template <typename L, typename I = intersect_t> static int cmd_step(char** a, int nmat, float threshold = 0)
{
	auto tris = read_tris(a[0]);
	const uint32_t seed = (uint32_t)std::stoul(a[2]);
	auto states = read_vec<St>(a[3]);
	std::vector<material<L>> mats;
	for (int i = 0; i < nmat; ++i)
		mats.push_back(material<L>(nbl::hdf5_file(a[5 + i])));
	geometry_t g = geometry_t::create(tris);
	I inter;
	Out o(a[4]);
	for (size_t i = 0; i < states.size(); ++i) {
		const St& s = states[i];
		nbl::drivers::simple_cpu_driver<L, I, geometry_t> d(g, inter, mats, threshold, seed + (uint32_t)i);
		particle p{ { s.pos[0], s.pos[1], s.pos[2] }, { s.dir[0], s.dir[1], s.dir[2] }, s.E };
		uint32_t tag = 0;
		d.push(&p, &tag, 1);
		auto& ps = d._particles.particles[0];
		ps.current_material = s.mat;
		ps.last_triangle = s.last_tri >= 0 ? g._triangles + s.last_tri : nullptr;
		d.init(0);
		d.intersect(0);
		d.scatter(0);
		const auto& q = d._particles.particles[0];
		o.put((int32_t)q.status);
		o.put((int32_t)q.next_scatter);
		o.put((int32_t)q.current_material);
		o.put((int32_t)(q.last_triangle ? q.last_triangle - g._triangles : -1));
		o.put(q.particle_data);
		const int32_t nse = (int32_t)d._particles.particles.size() - 1;
		o.put(nse);
		for (int32_t k = 1; k <= nse; ++k) {
			o.put(d._particles.particles[k].particle_data);
			o.put((int32_t)d._particles.particles[k].current_material);
		}
	}
	return 0;
}

// pri.bin: .pri records.  Output: detected records (float pos[3], dir[3], E,
// uint32 tag) in the driver's flush order; then a trailer int64 count.
template <typename L, typename I = intersect_t> static int cmd_simulate(char** a, int nmat, float threshold = 0)
{
	auto tris = read_tris(a[0]);
	const uint32_t seed = (uint32_t)std::stoul(a[2]);
	std::vector<material<L>> mats;
	for (int i = 0; i < nmat; ++i)
		mats.push_back(material<L>(nbl::hdf5_file(a[5 + i])));
	geometry_t g = geometry_t::create(tris);
	std::vector<particle> prim;
	std::vector<int2> pix;
	std::tie(prim, pix) = nbl::load_pri_file(a[3], g.AABB_min(), g.AABB_max(), 1e9f);
	I inter;
	nbl::drivers::simple_cpu_driver<L, I, geometry_t> d(g, inter, mats, threshold, seed);
	Out o(a[4]);
	int64_t n = 0;
	for (size_t i = 0; i < prim.size(); ++i) {
		uint32_t tag = (uint32_t)i;
		d.push(&prim[i], &tag, 1);
		d.simulate_to_end();
		d.flush_detected([&](particle p, uint32_t t) {
			o.put(p);
			o.put(t);
			++n;
		});
	}
	o.put(n);
	return 0;
}

template <typename L> static int cmd_edep(char** a, int nmat)
{
	auto tris = read_tris(a[0]);
	const uint32_t seed = (uint32_t)std::stoul(a[2]);
	std::vector<material<L>> mats;
	for (int i = 0; i < nmat; ++i)
		mats.push_back(material<L>(nbl::hdf5_file(a[5 + i])));
	geometry_t g = geometry_t::create(tris);
	std::vector<particle> prim;
	std::vector<int2> pix;
	std::tie(prim, pix) = nbl::load_pri_file(a[3], g.AABB_min(), g.AABB_max(), 1e9f);
	intersect_t inter;
	nbl::drivers::energydeposit_cpu_driver<L, intersect_t, geometry_t> d(g, inter, mats, 0, seed);
	struct Dep { float pos[3], E, dE; uint32_t tag; };
	std::vector<Dep> deps;
	std::vector<std::pair<particle, uint32_t>> det;
	for (size_t i = 0; i < prim.size(); ++i) {
		uint32_t tag = (uint32_t)i;
		d.push(&prim[i], &tag, 1);
		d.simulate_to_end([&](particle const& before, particle const& after, uint32_t t) {
			deps.push_back({ { after.pos.x, after.pos.y, after.pos.z }, before.kin_energy,
				before.kin_energy - after.kin_energy, t });
		});
		d.flush_detected([&](particle p, uint32_t t) { det.push_back({ p, t }); });
	}
	Out o(a[4]);
	for (const Dep& r : deps)
		o.put(r);
	o.put((int64_t)deps.size());
	for (const auto& r : det) {
		o.put(r.first);
		o.put(r.second);
	}
	o.put((int64_t)det.size());
	return 0;
}

// stepx / simulatex: argv a = <tris> <variant> <threshold> <seed> ...; the
// variant and threshold are taken out so that the rest reads as for step /
// simulate.
static int cmd_variant(bool sim, char** a, int na)
{
	const std::string v = a[1];
	const float threshold = std::stof(a[2]);
	char* b[64];
	const int nb = na - 1;
	if (nb > 64)
		throw std::runtime_error("too many arguments");
	b[0] = a[0];
	b[1] = a[1];
	for (int i = 2; i < nb; ++i)
		b[i] = a[i + 1];
	const int nmat = nb - 5;
	if (v == "penn")
		return sim ? cmd_simulate<penn_list>(b, nmat, threshold) : cmd_step<penn_list>(b, nmat, threshold);
	if (v == "kieft")
		return sim ? cmd_simulate<kieft_list>(b, nmat, threshold) : cmd_step<kieft_list>(b, nmat, threshold);
	if (v == "penn_abs")
		return sim ? cmd_simulate<penn_list, intersect_abs>(b, nmat, threshold)
			: cmd_step<penn_list, intersect_abs>(b, nmat, threshold);
	if (v == "penn_alt")
		return sim ? cmd_simulate<penn_alt_list, intersect_penn_alt>(b, nmat, threshold)
			: cmd_step<penn_alt_list, intersect_penn_alt>(b, nmat, threshold);
	if (v == "kieft_alt")
		return sim ? cmd_simulate<kieft_alt_list, intersect_kieft_alt>(b, nmat, threshold)
			: cmd_step<kieft_alt_list, intersect_kieft_alt>(b, nmat, threshold);
	throw std::runtime_error("unknown variant " + v);
}

static int cmd_pri(char** a)
{
	const float big = std::numeric_limits<float>::max();
	std::vector<particle> prim;
	std::vector<int2> pix;
	std::tie(prim, pix) = nbl::load_pri_file(a[0], { -big, -big, -big }, { big, big, big }, big);
	if (std::stoi(a[2]))
		nbl::sort_pri_file(prim, pix);
	nbl::prescan_shuffle(prim, pix, (size_t)std::stoull(a[1]));
	Out o(a[3]);
	for (size_t i = 0; i < prim.size(); ++i) {
		o.put(prim[i]);
		o.put(pix[i]);
	}
	return 0;
}

int main(int argc, char** argv)
{
	if (argc < 2) {
		std::fprintf(stderr, "usage: harness <command> ...\n");
		return 2;
	}
	const std::string c = argv[1];
	char** a = argv + 2;
	const int na = argc - 2;
	try {
		if (c == "octree") return cmd_octree(a);
		if (c == "propagate") return cmd_propagate(a);
		if (c == "intersect") return cmd_intersect(a);
		if (c == "pri") return cmd_pri(a);
		if (c == "stepx") return cmd_variant(false, a, na);
		if (c == "simulatex") return cmd_variant(true, a, na);
		const bool penn = na > 1 && std::string(a[1]) == "penn";
		if (c == "material") return penn ? cmd_material<penn_list>(a) : cmd_material<kieft_list>(a);
		if (c == "probe") return penn ? cmd_probe<penn_list>(a) : cmd_probe<kieft_list>(a);
		if (c == "step") return penn ? cmd_step<penn_list>(a, na - 5) : cmd_step<kieft_list>(a, na - 5);
		if (c == "simulate") return penn ? cmd_simulate<penn_list>(a, na - 5) : cmd_simulate<kieft_list>(a, na - 5);
		if (c == "edep") return penn ? cmd_edep<penn_list>(a, na - 5) : cmd_edep<kieft_list>(a, na - 5);
	} catch (const std::exception& e) {
		std::fprintf(stderr, "harness error: %s\n", e.what());
		return 1;
	}
	std::fprintf(stderr, "unknown command %s\n", c.c_str());
	return 2;
}
