/*
 * nebula.h -- C ABI of the MI355X electron-tracking library (libnebula.so).
 *
 * Plain pointers and sizes only.  Every function returns an int status:
 * NBL_OK (0) or a negative error class; nbl_last_error() gives the message of
 * the calling thread's last failure.  No exception crosses this boundary.
 *
 * Reference interfaces replaced (paths relative to the reference's source/):
 * each declaration names the reference function or class it stands in for.
 */
#ifndef NEBULA_H
#define NEBULA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ errors ---- */
enum {
	NBL_OK = 0,
	NBL_ERR_ARG = -1,      /* invalid argument / malformed input           */
	NBL_ERR_IO = -2,       /* file cannot be opened / truncated            */
	NBL_ERR_FORMAT = -3,   /* wrong column count, missing dataset, ...     */
	NBL_ERR_ALLOC = -4,    /* host or device allocation failed             */
	NBL_ERR_DEVICE = -5,   /* HIP runtime error                            */
	NBL_ERR_CAPACITY = -6, /* request beyond a stated maximum              */
	NBL_ERR_UNSUPPORTED = -7
};

/* Message of the calling thread's last error (empty if none). */
const char* nbl_last_error(void);
/* Library version: major*10000 + minor*100 + patch. */
int nbl_version(void);

/* ------------------------------------------------------------- types ---- */
/* core/triangle.h:8-77 -- 44 bytes: materials, r0, e1 = r1-r0, e2 = r2-r0.
 * material_in is the side the normal e1 x e2 points away from. */
typedef struct {
	int32_t material_in, material_out;
	float r0[3], e1[3], e2[3];
} nbl_triangle44;

/* core/particle.h:5-10 -- 28 bytes: position (nm), direction, energy (eV). */
typedef struct {
	float pos[3], dir[3], kin_energy;
} nbl_particle28;

/* io/load_pri_file.cpp:75,110-121 -- one .pri record / one output record of
 * gpu_main.cu:216-223: 7 floats (pos, dir, E) + pixel (2 x int32). */
typedef struct {
	float pos[3], dir[3], kin_energy;
	int32_t pixel[2];
} nbl_record36;

/* One detected electron as drained from a driver: the particle, its primary
 * tag and the stream id of the electron (primary: its index; secondary: a
 * hash of its parent's stream id and step). */
typedef struct {
	float pos[3], dir[3], kin_energy;
	uint32_t tag;
	uint64_t sid;
} nbl_detected;

/* Special material indices (core/material_manager.h:77-86). */
enum {
	NBL_MAT_NOP = -128, NBL_MAT_TERMINATOR = -127, NBL_MAT_DETECTOR = -126,
	NBL_MAT_DETECTOR_LT50 = -125, NBL_MAT_DETECTOR_GE50 = -124,
	NBL_MAT_VACUUM = -123, NBL_MAT_MIRROR = -122
};

/* Inelastic models: full_penn (physics/full_penn.h, the reference default of
 * physics_config.h:21-24) or kieft_inelastic (physics/kieft/inelastic.h). */
enum { NBL_INEL_PENN = 0, NBL_INEL_KIEFT = 1 };

/* Physics flags: the template switches of physics_config.h:21-38 and of
 * kieft_inelastic<...> made runtime.  NBL_PHYS_DEFAULT matches the
 * reference's compiled configuration. */
enum {
	NBL_PHYS_EL_PHONON_LOSS = 1u << 0,        /* kieft_elastic acoustic_phonon_loss */
	NBL_PHYS_EL_ATOMIC_RECOIL = 1u << 1,      /* kieft_elastic atomic_recoil_loss   */
	NBL_PHYS_PN_GENERATE_SE = 1u << 2,        /* full_penn generate_secondary       */
	NBL_PHYS_KI_OPTICAL_PHONON = 1u << 3,     /* kieft_inelastic optical_phonon_loss */
	NBL_PHYS_KI_GENERATE_SE = 1u << 4,        /* kieft_inelastic generate_secondary */
	NBL_PHYS_KI_INST_MOMENTUM = 1u << 5,      /* kieft_inelastic instantaneous_momentum */
	NBL_PHYS_KI_MOMENTUM_CONS = 1u << 6,      /* kieft_inelastic momentum_conservation */
	NBL_PHYS_BI_QUANTUM_TRANSMISSION = 1u << 7, /* boundary_intersect<T,.,.> */
	NBL_PHYS_BI_REFRACTION = 1u << 8,         /* boundary_intersect<.,T,.>  */
	NBL_PHYS_BI_ABSORPTION = 1u << 9,         /* boundary_intersect<.,.,T>  */
	NBL_PHYS_DEFAULT = 0x1ffu & ~(1u << 9)
};

/* A raw dataset of a .mat file (io/hdf5_file.cpp:229-289): float data in
 * row-major order, the dataset unit in internal units (eV = nm = 1, the
 * reference's gram = 6.2415e15), and per axis the first and last dimension
 * scale value (times the scale's unit) with ax_present = 0 when no scale is
 * attached (a linear axis then spans [0, 1]). */
typedef struct {
	const float* data;
	double unit;
	int32_t ndim;
	int32_t dims[3];
	double ax_front[3], ax_back[3];
	int32_t ax_present[3];
} nbl_raw_dataset;

/* The contents of one .mat file.  Attributes in internal units. */
typedef struct {
	double barrier, fermi, band_gap, phonon_loss, effective_A;
	int32_t has_band_gap;
	int32_t inel_model;                 /* NBL_INEL_* */
	nbl_raw_dataset el_imfp, el_icdf;   /* /elastic/imfp, /elastic/costheta_icdf */
	nbl_raw_dataset in_imfp, in_icdf;   /* /full_penn/{imfp,omega_icdf} or
	                                       /inelastic_kieft/{imfp,w0_icdf}  */
	nbl_raw_dataset pn_q, pn_ion;       /* /full_penn/q_icdf, 3-D /ionization/binding_icdf */
	nbl_raw_dataset ki_ion;             /* 2-D /ionization/binding_icdf */
	const double* outer_shells;         /* /ionization/outer_shells values */
	int32_t n_outer_shells;
	double outer_shells_unit;
} nbl_raw_material;

/* One prepared table: float data and per-axis (min, step) after the
 * reference's create() preprocessing.  Axis index = (v - min) * step. */
typedef struct {
	float* data;
	int32_t ndim;
	int32_t dims[3];
	float min[3], step[3];
} nbl_table;

/* A material after preprocessing (core/material.h:19-46 with the members of
 * kieft_elastic, full_penn and kieft_inelastic). */
typedef struct {
	float barrier, phonon_loss, recoil_const, fermi, band_gap;
	int32_t inel_model;
	nbl_table el_imfp, el_icdf, in_imfp, in_icdf, pn_q, pn_ion, ki_ion;
} nbl_material_host;

typedef struct nbl_device nbl_device;
typedef struct nbl_geometry nbl_geometry;
typedef struct nbl_material nbl_material;
typedef struct nbl_driver nbl_driver;

/* ------------------------------------------------- host-side (no GPU) ---- */
/* Free memory returned by this library. */
void nbl_free(void* p);

/* nbl::load_tri_file (io/load_tri_file.cpp:8-60).  A missing file gives
 * NBL_ERR_IO; a line without 11 columns gives NBL_ERR_FORMAT with the
 * reference's message. */
int nbl_load_tri_file(const char* path, nbl_triangle44** tris, int32_t* n);

/* nbl::load_pri_file (io/load_pri_file.cpp:72-184): reads 36-byte records;
 * a truncated file gives NBL_ERR_FORMAT ("Unexpected end of primaries
 * file").  The domain / energy / direction checks only count warnings:
 * warnings[0..3] = outside geometry, low energy, above max_E, bad direction. */
int nbl_load_pri_file(const char* path, const float aabb_min[3], const float aabb_max[3],
	float max_E, nbl_record36** records, int64_t* n, int64_t warnings[4]);

/* nbl::prescan_shuffle (io/load_pri_file.cpp:231-244): partial Fisher-Yates of
 * the first prescan_size records with a default-seeded minstd_rand0 engine. */
int nbl_prescan_shuffle(nbl_record36* records, int64_t n, int64_t prescan_size);

/* nbl::sort_pri_file (io/load_pri_file.cpp:186-229): reorder records by the
 * Morton code of their position over the primaries' bounding box (the
 * nebula_gpu --sort-primaries option). */
int nbl_sort_primaries(nbl_record36* records, int64_t n);

/* octree_builder::octree_root + linearized_octree
 * (geometry/octree/octree_builder.cpp:171-352).  Outputs are allocated by the
 * library; release with nbl_free. */
int nbl_octree_build(const nbl_triangle44* tris, int32_t ntri,
	int32_t** octree_data, int32_t* n_octree,
	nbl_triangle44** sorted_tris, int32_t* n_sorted,
	float center[3], float halfsize[3]);

/* material<...>::material(hdf5_file) -- the create() of kieft_elastic,
 * full_penn / kieft_inelastic (physics/kieft/elastic.h:120-166,
 * physics/full_penn.h:294-406, physics/kieft/inelastic.h:212-327). */
int nbl_material_prepare(const nbl_raw_material* raw, nbl_material_host** out);
void nbl_material_host_free(nbl_material_host* m);
/* scatter_list::get_max_energy (core/scatter_list.h:59-64). */
float nbl_material_max_energy(const nbl_material_host* m);

/* hdf5_file (io/hdf5_file.cpp) reading a .mat file into a prepared material.
 * Needs libnbl_hdf5.so (HDF5 C library); NBL_ERR_UNSUPPORTED if absent. */
int nbl_material_load_file(const char* path, int32_t inel_model, nbl_material_host** out);

/* ------------------------------------------------------------ device ---- */
/* cudaSetDevice + the allocations of gpu_main.cu:92-123. */
int nbl_device_open(int ordinal, nbl_device** dev);
int nbl_device_close(nbl_device* dev);
int nbl_device_synchronize(nbl_device* dev);

/* octree<true>::create (geometry/octree.inl:282-303): upload a linearised
 * octree (layout of octree_builder::linearized_octree).  Device memory:
 * 4 B (table) + 8 B (entries with their leaf-list lengths) + 32 B
 * (screening records) per table entry and 48 B per triangle.  Material ids must fit in 16 bits
 * (NBL_ERR_ARG otherwise). */
int nbl_geometry_upload(nbl_device* dev, const int32_t* octree_data, int64_t n_octree,
	const nbl_triangle44* tris, int64_t ntri, const float center[3], const float halfsize[3],
	nbl_geometry** out);
int nbl_geometry_destroy(nbl_geometry* g);

/* material<gpu>(material<cpu>) clone (core/material.inl:20-30). */
int nbl_material_upload(nbl_device* dev, const nbl_material_host* m, nbl_material** out);
int nbl_material_destroy(nbl_material* m);

/* gpu_driver ctor (drivers/gpu/gpu_driver.inl:51-89).  capacity = target
 * number of particles in flight; steps_per_launch = steps each particle takes
 * per kernel launch.  The driver keeps its own copies of geometry and
 * material descriptors (the objects must outlive it). */
int nbl_driver_create(nbl_device* dev, uint32_t capacity, const nbl_geometry* geom,
	const nbl_material* const* mats, int32_t n_mats, uint32_t physics_flags,
	float energy_threshold, uint64_t seed, uint32_t steps_per_launch, nbl_driver** out);
int nbl_driver_destroy(nbl_driver* d);

/* gpu_driver::push / push_to_buffer + push_to_simulation
 * (drivers/gpu/gpu_driver.inl:126-138,164-253): queue primaries on the device.
 * Primary i of this call gets tag tags[i] (or running index if tags is NULL)
 * and stream id sid_base + i. */
int nbl_driver_push(nbl_driver* d, const nbl_particle28* particles, const uint32_t* tags,
	uint32_t n, uint64_t sid_base, uint32_t* pushed);

/* Several scans in one driver (no counterpart: the reference runs one
 * geometry per process).  add_world registers another geometry with its own
 * material list (world 0 is the creation geometry) and returns its id;
 * push_world queues primaries that track in that world.  Secondaries stay
 * in their parent's world.  Physics flags, threshold and seed are shared;
 * nbl_driver_set_material_params applies to the material of that index in
 * every world that has one. */
int nbl_driver_add_world(nbl_driver* d, const nbl_geometry* geom, const nbl_material* const* mats,
	int32_t n_mats, int32_t* world);
int nbl_driver_push_world(nbl_driver* d, const nbl_particle28* particles, const uint32_t* tags,
	uint32_t n, uint64_t sid_base, int32_t world, uint32_t* pushed);

/* frame_size x gpu_driver::do_iteration (gpu_main.cu:209-210,
 * drivers/gpu/gpu_driver.inl:313-346).  Enqueues n_iter iterations on the
 * driver's stream without waiting. */
int nbl_driver_iterate(nbl_driver* d, uint32_t n_iter);

/* Energy deposits (drivers/cpu/energydeposit_cpu_driver.h:44-93): one
 * record per scattering event (elastic or inelastic, not a boundary) that
 * changes the electron's energy: position after the event, energy before
 * it, energy lost, tag of the primary, stream id of the electron.  Enable
 * before tracking; every iteration can write up to (K + 1) records per
 * particle in flight, so drain after every nbl_driver_iterate call (a drain
 * that finds records were dropped returns NBL_ERR_CAPACITY). */
typedef struct {
	float pos[3];
	float energy;
	float deposit;
	uint32_t tag;
	uint64_t sid;
} nbl_deposit;
int nbl_driver_set_deposits(nbl_driver* d, int enable);
int nbl_driver_drain_deposits(nbl_driver* d, nbl_deposit* out, uint64_t cap, uint64_t* n);

/* Launch configuration: steps each particle takes per k_step launch (1 to
 * 1024), k_step blocks (256 threads) per compute unit (1 to 128; default 64,
 * and never more than the capacity needs), and particles in flight to aim
 * for (<= capacity; the reference's batch sizing, gpu_main.cu:167-175). */
int nbl_driver_set_launch(nbl_driver* d, uint32_t steps_per_launch, uint32_t blocks_per_cu,
	uint32_t target);

/* Replay each nbl_driver_iterate(d, n > 1) from a captured hipGraph of n
 * iterations (captured once per n and launch configuration). */
int nbl_driver_set_graph(nbl_driver* d, int enable);

/* The prescan autotune of gpu_main.cu:136-175 made a measured search:
 * coordinate search over steps per launch {4,8,16,32}, blocks per compute
 * unit {8,16,32,64} and particles in flight {capacity/4, /2, 1}, each trial
 * `iters_per_trial` iterations of the queued primaries (real work: their
 * detected electrons stay in the buffer for the next drain).  Leaves the
 * fastest configuration (particle-steps/s) set on the driver. */
typedef struct {
	uint32_t steps_per_launch, blocks_per_cu, target, trials;
	double steps_per_s;
} nbl_tune_result;
int nbl_driver_autotune(nbl_driver* d, uint32_t iters_per_trial, nbl_tune_result* out);

/* gpu_driver::get_running_count / get_detected_count
 * (drivers/gpu/gpu_driver.inl:140-162).  Waits for the stream.  *queued =
 * primaries pushed but not yet moved into the particle buffer. */
int nbl_driver_counts(nbl_driver* d, uint32_t* running, uint32_t* detected, uint64_t* queued);

/* buffer_detected + flush_buffered (drivers/gpu/gpu_driver.inl:255-307):
 * copy out up to cap detected electrons and remove them from the device. */
int nbl_driver_drain_detected(nbl_driver* d, nbl_detected* out, uint32_t cap, uint32_t* n);

/* Total particle-steps taken so far (init calls on active particles). */
int nbl_driver_steps(nbl_driver* d, uint64_t* steps);

/* Run until every pushed primary and its cascade has finished, draining
 * detected electrons into out (cap records).  Returns NBL_ERR_CAPACITY if
 * more than cap electrons were detected. */
int nbl_driver_run(nbl_driver* d, nbl_detected* out, uint64_t cap, uint64_t* n_detected);

/* The frame loop of gpu_main.cu:200-230 with buffer_detected/flush_buffered
 * (drivers/gpu/gpu_driver.inl:255-307): run until every pushed primary and
 * its cascade has finished, frame_iters iterations per frame.  After each
 * frame its detected electrons are copied on the stream into pinned memory
 * and the next frame is enqueued at once; then sink(user, records, n,
 * running, queued) receives them (running / queued: particles in flight
 * and primaries not yet started, for progress output) while the GPU runs
 * that frame.  The records are valid until the sink returns; a nonzero
 * return stops the run (NBL_ERR_ARG). */
typedef int (*nbl_record_sink)(void* user, const nbl_detected* records, uint32_t n, uint32_t running,
	uint64_t queued);
int nbl_driver_run_frames(nbl_driver* d, uint32_t frame_iters, nbl_record_sink sink, void* user);

/* The HIP stream the driver launches on (a hipStream_t). */
void* nbl_driver_stream(nbl_driver* d);

/* Tracking without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code between a graph's capture calls,
 * serving loops, harnesses that forbid hipMalloc).  The same tracking as
 * nbl_driver_create + nbl_driver_push + nbl_driver_run with the plain
 * kernels: primaries[i] (DEVICE array, n entries) gets tag tags[i] (device
 * array, or NULL: tag i) and stream id sid_base + i; every detected electron
 * of the cascades is written to `out` (DEVICE array of out_cap records, in
 * detection order); *n_detected and *steps (particle-steps) are host
 * outputs.  The library allocates nothing: particle buffers, inelastic
 * queue, spill stack, control block and material descriptors live in one
 * fixed device array of 49 MiB (51 398 656 bytes) per GPU, held until the process ends and
 * shared by every call on that GPU (calls that overlap on one GPU race).
 * Limits: capacity (particles in flight) in [64, 65536], at most 64
 * materials (NBL_ERR_CAPACITY otherwise); the spill stack holds 2^19
 * secondaries and cannot grow, so a run that fills it with particles
 * waiting for secondary slots ends with NBL_ERR_CAPACITY, as does one that
 * detects more than out_cap electrons (*n_detected = out_cap then).
 * Launches on the device's stream and returns when the run has ended.
 * No scores, deposits or several worlds. */
int nbl_track_noalloc(nbl_device* dev, const nbl_geometry* geom, const nbl_material* const* mats,
	int32_t n_mats, uint32_t physics_flags, float energy_threshold, uint64_t seed, uint32_t capacity,
	uint32_t steps_per_launch, const nbl_particle28* primaries, const uint32_t* tags, uint32_t n,
	uint64_t sid_base, nbl_detected* out, uint32_t out_cap, uint32_t* n_detected, uint64_t* steps);

/* Introspection (no reference counterpart): the sizes of a driver's device
 * buffers, for tests and capacity planning.  A push appends behind the
 * queue's entries when it fits; otherwise only the entries not yet moved
 * into the particle buffer are kept, moved to the front (a drained queue
 * starts again from 0), and the queue grows only when those and the new
 * primaries exceed its capacity.  The spill stack of secondaries
 * that found no room doubles whenever it is found more than half full. */
typedef struct {
	uint64_t queue_capacity;    /* primaries the device queue can hold     */
	uint64_t queue_entries;     /* entries held (unread + read since reuse) */
	uint32_t slots;             /* particle-buffer entries (2 x capacity)  */
	uint32_t spill_capacity;    /* secondaries the spill stack can hold    */
	uint32_t spill_growths;     /* doublings of the spill stack so far     */
	uint32_t detected_capacity; /* detected records held between drains    */
} nbl_driver_info;
int nbl_driver_get_info(nbl_driver* d, nbl_driver_info* out);

/* Measurement: time every tracking-kernel launch with HIP events on the
 * driver's stream (no reference counterpart; replaces time_log of
 * common/time_log.h for the kernel).  nbl_driver_timing waits for the stream
 * and returns the summed kernel time and the number of launches since
 * timing was enabled. */
int nbl_driver_set_timing(nbl_driver* d, int enable);
int nbl_driver_timing(nbl_driver* d, double* step_ms_total, uint64_t* step_launches);

/* Page-locked host memory on `dev`'s runtime (hipHostMalloc).  A push from
 * it, or a drain into it, is one DMA at the link's rate instead of the
 * runtime's copy through bounce buffers.  The reference stages its pushes
 * and drains through cudaMallocHost buffers of its own
 * (drivers/gpu/gpu_driver.inl:81-83, 208-210); here the caller writes its
 * particles straight into such memory.  *ptr is null for bytes == 0.
 * Release with nbl_host_free. */
int nbl_host_alloc(nbl_device* dev, uint64_t bytes, void** ptr);
int nbl_host_free(void* ptr);

/* Measurement aid (no reference counterpart): free and total device memory
 * of `dev` (hipMemGetInfo), for reporting a run's footprint. */
int nbl_device_memory(nbl_device* dev, uint64_t* free_bytes, uint64_t* total_bytes);

/* Measurement aid (no reference counterpart): copy `bytes` between two
 * scratch device buffers with a plain copy kernel (k_copy) and return its
 * time; a known-size reference point for HBM traffic counters. */
int nbl_device_copy_probe(nbl_device* dev, uint64_t bytes, float* ms);

/* ------------------------------------------------------- test hooks ---- */
/* Batched octree::propagate (geometry/octree.inl:32-192): ray i from pos[3i]
 * along dir[3i] (already normalised) over dist[i], ignoring triangle
 * ignore_tri[i] (-1: none) and destination material ignore_mat[i].  Writes
 * the hit triangle index (-1: none) and distance.  Host arrays. */
int nbl_geometry_query(const nbl_geometry* g, const float* pos, const float* dir,
	const float* dist, const int32_t* ignore_tri, const int32_t* ignore_mat,
	int32_t* out_tri, float* out_dist, uint32_t n);

/* Evaluate table `which` (0 el_imfp, 1 el_icdf, 2 in_imfp, 3 in_icdf, 4 pn_q,
 * 5 pn_ion, 6 ki_ion) of an uploaded material at n points (x, y, z host
 * arrays; unused axes ignored).  mode 0 = get (interpolate), 1 =
 * get_rounddown, 2 = get_nearest (2-D only).
 * (common/util/table_{1,2,3}D.inl) */
int nbl_table_probe(const nbl_material* m, int32_t which, int32_t mode, const float* x,
	const float* y, const float* z, float* out, uint32_t n);

/* One step of particle i of a host array of states on the GPU: for parity
 * tests of the event physics.  state layout = nbl_step_state. */
typedef struct {
	float pos[3], dir[3], kin_energy;
	int32_t mat, last_tri;
	uint32_t tag, step, flags;
	uint64_t sid;
} nbl_step_state;
/* Applies one step (init + event) to each state with the driver's physics;
 * writes the new state, the event (0 none, 1 inelastic, 2 elastic,
 * 3 intersect), the outcome (0 alive, 1 terminated, 2 detected) and the
 * secondary electron made (se_made[i] = 1). */
int nbl_driver_step_states(nbl_driver* d, nbl_step_state* states, int32_t* event,
	int32_t* outcome, nbl_step_state* se, int32_t* se_made, uint32_t n);

/* ------------------------------------------ derivatives (no counterpart) --
 * The reference has no derivative of its outputs.  These entry points let a
 * caller estimate d(yield)/d(material parameter) by the likelihood-ratio
 * method: a run with scores enabled replays the same trajectories (same
 * seed and stream ids give the same random draws) and attaches to every
 * detected electron the FP64 score sum_steps d log p(step) / d theta over
 * its own steps and its ancestors' steps up to its creation.  The mean over
 * primaries of (yield indicator x score) is the derivative of the mean
 * yield.  Parameters per material: kind 0 = elastic mean-free-path scale,
 * 1 = inelastic mean-free-path scale, 2 = barrier (work function + Fermi
 * energy, eV).  Scores cover the random choices (free-path race, boundary
 * transmission / absorption); the barrier's deterministic effects (energy
 * shift, refraction, threshold) are not part of them. */
#define NBL_KIND_EL_MFP_SCALE 0
#define NBL_KIND_IN_MFP_SCALE 1
#define NBL_KIND_BARRIER 2
#define NBL_MAX_SCORE_PARAMS 8

/* Model parameters of material `material` for this driver (in every world
 * that has a material of that index): free paths are sampled with the mean
 * free path x scale; `barrier` replaces the file's. */
int nbl_driver_set_material_params(nbl_driver* d, int32_t material, float el_mfp_scale,
	float in_mfp_scale, float barrier);
int nbl_driver_get_material_params(nbl_driver* d, int32_t material, float* el_mfp_scale,
	float* in_mfp_scale, float* barrier);

/* Enable scores for n (<= 8) parameters (materials[i], kinds[i]), or
 * disable with n = 0.  Only while nothing is in flight, queued or waiting
 * to be drained (NBL_ERR_ARG otherwise); primaries start with score 0. */
int nbl_driver_set_score_params(nbl_driver* d, const int32_t* materials, const int32_t* kinds,
	int32_t n);

/* nbl_driver_drain_detected / nbl_driver_run plus scores[i * n + k] for
 * record i and parameter k. */
int nbl_driver_drain_detected_scores(nbl_driver* d, nbl_detected* out, double* scores,
	uint32_t cap, uint32_t* n);
int nbl_driver_run_scores(nbl_driver* d, nbl_detected* out, double* scores, uint64_t cap,
	uint64_t* n_detected);

#ifdef __cplusplus
}
#endif
#endif /* NEBULA_H */
