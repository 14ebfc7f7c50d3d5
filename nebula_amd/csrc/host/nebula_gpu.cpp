// nebula_gpu: command-line driver over the C ABI, a drop-in for the
// reference's nebula_gpu (gpu_main.cu:92-445).
//
//   nebula_gpu [options] <geometry.tri> <primaries.pri> [material0.mat] .. [materialN.mat]
//
// Same options and defaults (gpu_main.cu:252-258), same `--key=value` /
// `--key value` syntax (common/cli_params.h:54-88), same stdout: one 36-byte
// record per detected electron, float32 pos[3], dir[3], energy, then int32
// pixel[2] of its primary (gpu_main.cu:214-223).  Diagnostics go to stderr.
//
// Sequence as in the reference: load and check the geometry, build the
// octree, load materials, load primaries (domain and energy warnings),
// optional Morton sort, prescan shuffle, prescan (frame and batch size from
// the running/detected curve, gpu_main.cu:136-175), then the frame loop.
//
// What differs, by design of the device side:
//  * All primaries are queued in HBM once (40 bytes each; 1e8 primaries take
//    4 GB of the 288 GB) and every iteration refills free slots from that
//    queue on the device.  The reference instead refills a host-side batch
//    of batch_size primaries once per frame.  batch_size is still computed
//    and reported; it has no effect here.
//  * frame_size sets how many iterations run between two reads of the
//    detected-electron buffer.  While a frame runs, the host writes the
//    previous frame's records.
//  * One GPU (the device given by --device, default 0).
//  * Two options the reference fixes at compile time: --inelastic
//    (penn = full_penn, the reference default, physics_config.h:21-24; or
//    kieft = kieft_inelastic) and --steps-per-launch.
//  * Energy deposits (the outputs of the reference's CPU-only
//    nebula_cpu_edep, cpu_energydep.cpp:170-204): --deposit-filename writes
//    one 28-byte record per scattering event that changes an electron's
//    energy (float32 position, energy before, energy lost; int32 pixel[2]),
//    --detect-filename moves the detected records off stdout.  Built as
//    nebula_edep, the binary takes nebula_cpu_edep's defaults
//    (detected.bin, deposits on stdout).
#include "nebula.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iomanip>
#include <iostream>
#include <map>
#include <mutex>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

// --key=value / --key value; anything else is positional.  Unknown keys are
// an error (common/cli_params.h:54-88).  Values are read with operator>>,
// so booleans are 0/1 as in the reference.
struct Options {
	struct Opt { std::string desc, value; };
	std::map<std::string, Opt> opts;
	std::vector<std::string> positional;
	std::string program;

	template <typename T> void add(const std::string& k, const std::string& d, const T& def)
	{
		std::stringstream s;
		s << def;
		opts[k] = { d, s.str() };
	}
	void parse(int argc, char** argv)
	{
		program = argv[0];
		for (int i = 1; i < argc; ++i) {
			const std::string item = argv[i];
			if (item.size() >= 2 && item[0] == '-' && item[1] == '-') {
				std::string key, value;
				const size_t split = item.find('=', 2);
				if (split == std::string::npos) {
					key = item.substr(2);
					value = (++i < argc) ? argv[i] : "";
					if (value == "=")
						value = (++i < argc) ? argv[i] : "";
				} else {
					key = item.substr(2, split - 2);
					value = item.substr(split + 1);
				}
				auto it = opts.find(key);
				if (it == opts.end())
					throw std::runtime_error("Unexpected key '" + key + "'");
				it->second.value = value;
			} else {
				positional.push_back(item);
			}
		}
	}
	template <typename T> T get(const std::string& k) const
	{
		T r{};
		std::stringstream(opts.at(k).value) >> r;
		return r;
	}
	void usage(std::ostream& o, const std::string& params) const
	{
		o << "Usage: " << program << " " << params << "\nOptions:\n";
		for (const auto& kv : opts)
			o << "\t--" << kv.first << ": " << kv.second.desc << " [" << kv.second.value << "]\n";
	}
};

// Named wall-clock phases printed at the end (common/time_log.h).
struct TimeLog {
	using clk = std::chrono::steady_clock;
	clk::time_point t0;
	std::vector<std::pair<std::string, clk::duration>> rows;
	void start() { t0 = clk::now(); }
	void stop(const std::string& name) { rows.push_back({ name, clk::now() - t0 }); }
	void print(std::ostream& o) const
	{
		size_t w = 0;
		for (const auto& r : rows)
			w = std::max(w, r.first.size());
		o << "Timing (seconds):\n";
		for (const auto& r : rows) {
			const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(r.second);
			o << std::setw((int)w) << r.first << ": " << std::fixed << std::setprecision(3)
			  << (ms.count() / 1000.) << '\n';
		}
	}
};

struct Failure : std::runtime_error {
	using std::runtime_error::runtime_error;
};

void check(int rc, const char* what)
{
	if (rc != NBL_OK)
		throw Failure(std::string(what) + ": " + nbl_last_error());
}

const char* kParams = "[options] <geometry.tri> <primaries.pri> [material0.mat] .. [materialN.mat]";

FILE* open_output(const std::string& name)
{
	if (name == "stdout")
		return stdout;
	FILE* f = std::fopen(name.c_str(), "wb");
	if (!f)
		throw Failure("cannot open " + name + " for writing");
	return f;
}

// Energy deposits -> 28-byte records (cpu_energydep.cpp:192-199).
void write_deposits(const std::vector<nbl_deposit>& dep, uint64_t n,
	const std::vector<nbl_record36>& pri, FILE* out, std::vector<char>& buf)
{
	struct Rec { float pos[3], energy, deposit; int32_t pixel[2]; };
	static_assert(sizeof(Rec) == 28, "record layout");
	buf.resize((size_t)n * sizeof(Rec));
	for (uint64_t i = 0; i < n; ++i) {
		Rec r;
		std::memcpy(r.pos, dep[i].pos, sizeof r.pos);
		r.energy = dep[i].energy;
		r.deposit = dep[i].deposit;
		r.pixel[0] = pri[dep[i].tag].pixel[0];
		r.pixel[1] = pri[dep[i].tag].pixel[1];
		std::memcpy(buf.data() + (size_t)i * sizeof r, &r, sizeof r);
	}
	if (n && std::fwrite(buf.data(), sizeof(Rec), (size_t)n, out) != (size_t)n)
		throw Failure("writing energy deposits failed");
}

// Detected electrons -> 36-byte records in `buf`, pixel looked up by tag.
// The lookups are random reads into the primaries, so a large frame is
// split over several threads.
void pack_records(const nbl_detected* det, uint32_t n, const std::vector<nbl_record36>& pri,
	std::vector<char>& buf)
{
	buf.resize((size_t)n * sizeof(nbl_record36));
	auto pack = [&](uint32_t lo, uint32_t hi) {
		for (uint32_t i = lo; i < hi; ++i) {
			nbl_record36 r;
			std::memcpy(r.pos, det[i].pos, sizeof r.pos);
			std::memcpy(r.dir, det[i].dir, sizeof r.dir);
			r.kin_energy = det[i].kin_energy;
			r.pixel[0] = pri[det[i].tag].pixel[0];
			r.pixel[1] = pri[det[i].tag].pixel[1];
			std::memcpy(buf.data() + (size_t)i * sizeof r, &r, sizeof r);
		}
	};
	const uint32_t per_thread = 1u << 16;
	const uint32_t nt = std::min<uint32_t>({ (n + per_thread - 1) / per_thread, 16u,
		std::max(1u, std::thread::hardware_concurrency()) });
	if (nt <= 1) {
		pack(0, n);
		return;
	}
	std::vector<std::thread> th;
	const uint32_t chunk = (n + nt - 1) / nt;
	// the chunks without a thread (all but the first, if none could be
	// started) are packed here; started threads are always joined
	uint32_t started = 1;
	try {
		for (; started < nt; ++started)
			th.emplace_back(pack, std::min(n, started * chunk), std::min(n, (started + 1) * chunk));
	} catch (...) {
	}
	pack(0, std::min(n, chunk));
	pack(std::min(n, started * chunk), n);
	for (auto& x : th)
		x.join();
}

void write_records(const nbl_detected* det, uint32_t n,
	const std::vector<nbl_record36>& pri, FILE* out, std::vector<char>& buf)
{
	pack_records(det, n, pri, buf);
	if (n && std::fwrite(buf.data(), sizeof(nbl_record36), n, out) != n)
		throw Failure("writing detected electrons failed");
}

// Writes blocks of output records on a thread of its own, so that a slow
// reader of the output (a pipe) holds up the frame loop only once
// `max_pending` blocks wait to be written.  Written blocks are kept for
// reuse (at most `max_pending` of them).
class RecordWriter
{
public:
	RecordWriter(FILE* out, size_t max_pending) : out_(out), max_pending_(max_pending) {}
	~RecordWriter() { finish(); }

	// An empty block to fill: a written one if there is one.
	std::vector<char> take()
	{
		std::lock_guard<std::mutex> g(m_);
		if (spare_.empty())
			return {};
		std::vector<char> b = std::move(spare_.back());
		spare_.pop_back();
		return b;
	}

	// Queue a filled block; false once a write has failed.
	bool put(std::vector<char>&& b)
	{
		std::unique_lock<std::mutex> g(m_);
		room_.wait(g, [&] { return pending_.size() < max_pending_ || failed_; });
		if (failed_)
			return false;
		pending_.push_back(std::move(b));
		work_.notify_one();
		return true;
	}

	// Write everything queued and stop the thread; false if a write failed.
	bool finish()
	{
		{
			std::lock_guard<std::mutex> g(m_);
			done_ = true;
		}
		work_.notify_one();
		if (th_.joinable())
			th_.join();
		return !failed_;
	}

private:
	void loop()
	{
		std::unique_lock<std::mutex> g(m_);
		for (;;) {
			work_.wait(g, [&] { return !pending_.empty() || done_; });
			if (pending_.empty())
				return;
			std::vector<char> b = std::move(pending_.front());
			pending_.pop_front();
			if (!failed_) {
				g.unlock();
				const bool ok = std::fwrite(b.data(), 1, b.size(), out_) == b.size();
				g.lock();
				failed_ = failed_ || !ok;
			}
			if (spare_.size() < max_pending_) {
				b.clear();
				spare_.push_back(std::move(b));
			}
			room_.notify_all();
		}
	}

	FILE* out_;
	const size_t max_pending_;
	std::mutex m_;
	std::condition_variable work_, room_;
	std::deque<std::vector<char>> pending_;
	std::vector<std::vector<char>> spare_;
	bool done_ = false, failed_ = false;
	std::thread th_{ [this] { loop(); } };   // last: starts once the members above exist
};

int run(int argc, char** argv)
{
	std::clog << "This is nebula_amd (C ABI version " << nbl_version() << ")\n\n"
		"Physics models:\n"
		"  elastic:   kieft_elastic (tabulated IMFP and angular ICDF, phonon loss, recoil)\n"
		"  inelastic: full_penn (default) or kieft_inelastic (--inelastic)\n"
		"  boundary:  boundary_intersect (quantum-mechanical transmission, refraction)\n\n"
		<< std::string(80, '-') << "\n\n";

	Options p;
	p.add("energy-threshold", "Lowest energy to simulate", 0);
	p.add("capacity", "Electron capacity on the GPU", 1000000);
	p.add("prescan-size", "Number of electrons to use for prescan", 1000);
	p.add("batch-factor", "Multiplication factor for electron batch size", 0.9);
	p.add("seed", "Random seed", 0x78c7e39b);
	p.add("sort-primaries", "Sort primary electrons before simulation", false);
	p.add("inelastic", "Inelastic model: penn (full_penn) or kieft (kieft_inelastic)", "penn");
	p.add("steps-per-launch", "Steps each electron takes per kernel launch", 6);
	p.add("device", "GPU ordinal", 0);
#ifdef NBL_EDEP
	p.add("detect-filename", "Filename for detected electrons", "detected.bin");
	p.add("deposit-filename", "Filename for energy deposits", "stdout");
#else
	p.add("detect-filename", "Filename for detected electrons", "stdout");
	p.add("deposit-filename", "Filename for energy deposits (empty: none)", "");
#endif
	p.parse(argc, argv);
	const float energy_threshold = p.get<float>("energy-threshold");
	const uint32_t capacity = p.get<uint32_t>("capacity");
	const uint32_t prescan_size = p.get<uint32_t>("prescan-size");
	const float batch_factor = p.get<float>("batch-factor");
	const unsigned int seed = p.get<unsigned int>("seed");
	const bool sort_primaries = p.get<bool>("sort-primaries");
	const std::string inel = p.get<std::string>("inelastic");
	const uint32_t steps_per_launch = p.get<uint32_t>("steps-per-launch");
	const int device = p.get<int>("device");
	const std::string detect_name = p.opts.at("detect-filename").value;
	const std::string deposit_name = p.opts.at("deposit-filename").value;
	const bool deposits = !deposit_name.empty();

	TimeLog timer;
	const auto& pos = p.positional;
	if (pos.size() < 3 || capacity <= 0 || prescan_size <= 0 || !(batch_factor > 0) ||
		(inel != "penn" && inel != "kieft") || steps_per_launch == 0) {
		p.usage(std::clog, kParams);
		return 1;
	}
	const int32_t inel_model = inel == "penn" ? NBL_INEL_PENN : NBL_INEL_KIEFT;
	// The reference gives GPU i the i-th draw of mt19937(seed) (gpu_main.cu:274,293).
	std::mt19937 seeder(seed);
	const uint64_t driver_seed = seeder();

	// ---------------------------------------------------------- geometry
	std::clog << "Loading geometry..." << std::endl;
	timer.start();
	nbl_triangle44* tris = nullptr;
	int32_t ntri = 0;
	const int rc_tri = nbl_load_tri_file(pos[0].c_str(), &tris, &ntri);
	timer.stop("Loading triangles");
	if (rc_tri == NBL_ERR_FORMAT)
		throw Failure(nbl_last_error());
	if (rc_tri != NBL_OK || ntri == 0) {
		nbl_free(tris);
		std::clog << "Error: could not load triangles!\n";
		p.usage(std::clog, kParams);
		return 1;
	}
	{
		int max_material = -1;
		for (int32_t i = 0; i < ntri; ++i)
			max_material = std::max({ max_material, tris[i].material_in, tris[i].material_out });
		const int n_mat_files = (int)pos.size() - 2;
		if (max_material > n_mat_files - 1) {
			std::clog << "Error: not enough materials provided for this geometry!\n"
				"  Expected " << (max_material + 1) << " materials, " << n_mat_files << " provided.\n";
			nbl_free(tris);
			return 1;
		}
		if (max_material < n_mat_files - 1)
			std::clog << "Warning: too many materials provided for this geometry!\n"
				"  Expected " << (max_material + 1) << " materials, " << n_mat_files << " provided.\n";
	}
	timer.start();
	int32_t *octree = nullptr, n_octree = 0, n_sorted = 0;
	nbl_triangle44* sorted = nullptr;
	float center[3], halfsize[3];
	check(nbl_octree_build(tris, ntri, &octree, &n_octree, &sorted, &n_sorted, center, halfsize),
		"building the octree");
	nbl_free(tris);
	timer.stop("Building acceleration structure");
	const float aabb_min[3] = { center[0] - halfsize[0], center[1] - halfsize[1], center[2] - halfsize[2] };
	const float aabb_max[3] = { center[0] + halfsize[0], center[1] + halfsize[1], center[2] + halfsize[2] };

	// --------------------------------------------------------- materials
	std::clog << "Loading materials..." << std::endl;
	timer.start();
	std::vector<nbl_material_host*> hmats;
	float max_E = 0;
	for (size_t i = 2; i < pos.size(); ++i) {
		nbl_material_host* m = nullptr;
		check(nbl_material_load_file(pos[i].c_str(), inel_model, &m), pos[i].c_str());
		hmats.push_back(m);
		const float e = nbl_material_max_energy(m);
		max_E = (i == 2) ? e : std::min(max_E, e);
	}
	timer.stop("Loading materials");

	// --------------------------------------------------------- primaries
	std::clog << "Loading primary electrons..." << std::endl;
	timer.start();
	nbl_record36* rec = nullptr;
	int64_t n_pri = 0, warn[4] = { 0, 0, 0, 0 };
	const int rc_pri = nbl_load_pri_file(pos[1].c_str(), aabb_min, aabb_max, max_E, &rec, &n_pri, warn);
	if (rc_pri == NBL_ERR_FORMAT)
		throw Failure(nbl_last_error());
	std::vector<nbl_record36> pri;
	if (rc_pri == NBL_OK)
		pri.assign(rec, rec + n_pri);
	nbl_free(rec);
	if (warn[0])
		std::clog << "WARNING: " << warn[0] << " primary electrons starting outside geometry." << std::endl;
	if (warn[1])
		std::clog << "WARNING: " << warn[1] << " primary electrons starting with low kinetic energy." << std::endl;
	if (warn[2])
		std::clog << "WARNING: " << warn[2] << " primary electrons starting with kinetic energy higher "
			"than what is available in the material files. Limit is " << (max_E / 1000) << " keV." << std::endl;
	if (warn[3])
		std::clog << "WARNING: " << warn[3] << " primary electrons start with unphysical direction vector." << std::endl;
	if (sort_primaries && !pri.empty())
		check(nbl_sort_primaries(pri.data(), (int64_t)pri.size()), "sorting primaries");
	if (!pri.empty())
		check(nbl_prescan_shuffle(pri.data(), (int64_t)pri.size(), prescan_size), "prescan shuffle");
	timer.stop("Loading primary electrons");
	if (pri.empty()) {
		std::clog << "Error: could not load primary electrons!\n";
		p.usage(std::clog, kParams);
		return 1;
	}
	std::clog << "\nLoaded " << ntri << " triangles.\n"
		<< "  min = {" << aabb_min[0] << ", " << aabb_min[1] << ", " << aabb_min[2] << "}\n"
		<< "  max = {" << aabb_max[0] << ", " << aabb_max[1] << ", " << aabb_max[2] << "}\n"
		<< "Loaded " << pri.size() << " primaries.\n"
		<< "Loaded " << hmats.size() << " materials.\n\n" << std::flush;

	// ------------------------------------------------------------ device
	nbl_device* dev = nullptr;
	check(nbl_device_open(device, &dev), "opening the GPU");
	std::clog << "Using GPU " << device << "\n" << std::endl;

	// Primary i (after sort and shuffle) has tag i and stream id i (the
	// driver's running tags: the two pushes below are its first).  The
	// particles are written into page-locked memory, so that each push is
	// one DMA.
	const size_t N = pri.size();
	void* part_mem = nullptr;
	check(nbl_host_alloc(dev, sizeof(nbl_particle28) * N, &part_mem), "page-locked memory for the primaries");
	nbl_particle28* part = static_cast<nbl_particle28*>(part_mem);
	for (size_t i = 0; i < N; ++i) {
		std::memcpy(part[i].pos, pri[i].pos, sizeof part[i].pos);
		std::memcpy(part[i].dir, pri[i].dir, sizeof part[i].dir);
		part[i].kin_energy = pri[i].kin_energy;
	}
	nbl_geometry* geom = nullptr;
	check(nbl_geometry_upload(dev, octree, n_octree, sorted, n_sorted, center, halfsize, &geom),
		"uploading the geometry");
	nbl_free(octree);
	nbl_free(sorted);
	std::vector<nbl_material*> dmats;
	for (nbl_material_host* m : hmats) {
		nbl_material* dm = nullptr;
		check(nbl_material_upload(dev, m, &dm), "uploading a material");
		dmats.push_back(dm);
		nbl_material_host_free(m);
	}
	nbl_driver* drv = nullptr;
	check(nbl_driver_create(dev, capacity, geom, dmats.data(), (int32_t)dmats.size(), NBL_PHYS_DEFAULT,
		energy_threshold, driver_seed, steps_per_launch, &drv), "creating the driver");
	FILE* det_out = open_output(detect_name);
	FILE* dep_out = deposits ? open_output(deposit_name) : nullptr;
	std::vector<nbl_deposit> depbuf;
	std::vector<char> dbuf;
	// deposits are read after every iteration (the buffer holds one)
	auto drain_deposits = [&]() {
		uint64_t nd = 0;
		const int rc = nbl_driver_drain_deposits(drv, nullptr, 0, &nd);
		if (rc == NBL_OK)
			return;
		if (rc != NBL_ERR_CAPACITY || nd == 0)
			check(rc, "reading energy deposits");
		depbuf.resize(nd);
		check(nbl_driver_drain_deposits(drv, depbuf.data(), nd, &nd), "reading energy deposits");
		write_deposits(depbuf, nd, pri, dep_out, dbuf);
	};
	if (deposits)
		check(nbl_driver_set_deposits(drv, 1), "enabling energy deposits");

	// ----------------------------------------------------------- prescan
	// Prescan primaries are real work; their detections are written first.
	timer.start();
	const uint32_t n_prescan = (uint32_t)std::min<size_t>(prescan_size, N);
	std::vector<std::pair<uint32_t, uint32_t>> stats;
	{
		uint32_t pushed = 0;
		check(nbl_driver_push(drv, part, nullptr, n_prescan, 0, &pushed), "push");
		stats.push_back({ pushed, 0 });
	}
	uint64_t last_steps = ~0ull;
	int stalled = 0;
	while (stats.back().first > 0) {
		check(nbl_driver_iterate(drv, 1), "iterate");
		if (deposits)
			drain_deposits();
		uint32_t running = 0, detected = 0;
		uint64_t queued = 0, steps = 0;
		check(nbl_driver_counts(drv, &running, &detected, &queued), "counts");
		check(nbl_driver_steps(drv, &steps), "steps");
		stats.push_back({ running + (uint32_t)queued, detected });
		std::clog << " \rExecuting pre-scan | running: " << stats.back().first
			<< " | detected: " << stats.back().second;
		stalled = (steps == last_steps) ? stalled + 1 : 0;
		last_steps = steps;
		if (stalled >= 64)
			throw Failure("prescan makes no progress: raise --capacity");
	}
	timer.stop("Prescan");
	const uint32_t frame_size = 1 + (uint32_t)std::distance(stats.begin(),
		std::max_element(stats.begin(), stats.end(),
			[](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
				return a.first < b.first; }));
	uint32_t batch_size;
	{
		float acc = 0;
		for (size_t i = 2 * (size_t)frame_size; i < stats.size(); i += frame_size)
			acc += stats[i].first / float(prescan_size);
		acc += 2 * stats[frame_size].first / float(prescan_size);
		acc += 2 * stats[frame_size].second / float(prescan_size);
		batch_size = acc > 0 ? uint32_t(batch_factor * capacity / acc) : capacity;
	}
	std::clog << "\nframe_size = " << frame_size << " | batch_size = " << batch_size << std::endl;

	// -------------------------------------------------------- frame loop
	timer.start();
	if (N > n_prescan) {
		uint32_t pushed = 0;
		check(nbl_driver_push(drv, part + n_prescan, nullptr, (uint32_t)(N - n_prescan), n_prescan,
			&pushed), "push");
	}
	std::vector<char> obuf;
	auto last_print = std::chrono::steady_clock::now();
	if (!deposits) {
		// The library's frame loop: each frame's records reach the sink
		// while the GPU already runs the next frame; the sink converts them
		// and hands them to the writer thread.
		RecordWriter writer(det_out, 4);
		struct Sink {
			const std::vector<nbl_record36>* pri;
			RecordWriter* writer;
			size_t n_total;
			std::chrono::steady_clock::time_point last_print;
			bool write_failed;
		} sk{ &pri, &writer, N, last_print, false };
		auto sink = [](void* user, const nbl_detected* recs, uint32_t n, uint32_t running, uint64_t queued) {
			Sink& k = *(Sink*)user;
			try {
				if (n) {
					std::vector<char> b = k.writer->take();
					pack_records(recs, n, *k.pri, b);
					if (!k.writer->put(std::move(b))) {
						k.write_failed = true;
						return 1;
					}
				}
			} catch (const std::exception&) {
				k.write_failed = true;
				return 1;
			}
			const auto now = std::chrono::steady_clock::now();
			const bool done = running == 0 && queued == 0;
			if (done || now - k.last_print >= std::chrono::seconds(1)) {
				std::clog << " \rProgress " << std::fixed << std::setprecision(2)
					<< 100.0 * (1.0 - (double)queued / (double)k.n_total) << "% | Running: " << running;
				k.last_print = now;
			}
			return 0;
		};
		const int rc = nbl_driver_run_frames(drv, frame_size, sink, &sk);
		if (!writer.finish())
			sk.write_failed = true;
		if (sk.write_failed)
			throw Failure("writing detected electrons failed");
		check(rc, "tracking");
	}
	std::vector<nbl_detected> det(deposits ? capacity : 0);
	last_steps = ~0ull;
	stalled = 0;
	while (deposits) {
		uint32_t got = 0, running = 0;
		uint64_t queued = 0, steps = 0;
		check(nbl_driver_drain_detected(drv, det.data(), capacity, &got), "reading detected electrons");
		check(nbl_driver_counts(drv, &running, nullptr, &queued), "counts");
		check(nbl_driver_steps(drv, &steps), "steps");
		const bool done = running == 0 && queued == 0;
		if (!done)
			check(nbl_driver_iterate(drv, deposits ? 1 : frame_size), "iterate");
		write_records(det.data(), got, pri, det_out, obuf);   // overlaps the frame just launched
		if (deposits && !done)
			drain_deposits();
		const auto now = std::chrono::steady_clock::now();
		if (done || now - last_print >= std::chrono::seconds(1)) {
			std::clog << " \rProgress " << std::fixed << std::setprecision(2)
				<< 100.0 * (1.0 - (double)queued / (double)N) << "% | Running: " << running;
			last_print = now;
		}
		if (done)
			break;
		stalled = (steps == last_steps && got == 0) ? stalled + 1 : 0;
		last_steps = steps;
		if (stalled >= 4)
			throw Failure("no progress: particle buffer and spill stack full of particles waiting "
				"for secondary-electron slots; raise --capacity");
	}
	std::fflush(det_out);
	if (det_out != stdout)
		std::fclose(det_out);
	if (dep_out) {
		std::fflush(dep_out);
		if (dep_out != stdout)
			std::fclose(dep_out);
	}
	timer.stop("Simulation");
	const int rc_sync = nbl_device_synchronize(dev);
	std::clog << (rc_sync == NBL_OK ? "\nSimulation successful!\n\n" : "\nSimulation ended with a GPU error\n\n");
	timer.print(std::clog);

	nbl_driver_destroy(drv);
	nbl_host_free(part_mem);
	for (nbl_material* m : dmats)
		nbl_material_destroy(m);
	nbl_geometry_destroy(geom);
	nbl_device_close(dev);
	return rc_sync == NBL_OK ? 0 : 1;
}

} // namespace

int main(int argc, char** argv)
{
	try {
		return run(argc, argv);
	} catch (const std::exception& e) {
		std::clog << "\nError: " << e.what() << std::endl;
		return 1;
	}
}
