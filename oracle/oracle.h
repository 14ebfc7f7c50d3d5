/*
 * oracle.h -- CPU restatement of Nebula's electron-tracking path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * nebula_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  Nothing in the product links, imports or calls it.
 *
 * Every function restates one reference function (file:line given beside it,
 * paths relative to the reference's source/ directory) with the same
 * floating-point operation order, so that, compiled with -ffp-contract=off on
 * x86-64, it is bit-identical to the reference's CPU build.
 *
 * Two builds are made from oracle.c:
 *   liboracle_ref.so  transcendentals are the C library's float functions, as
 *                     std::log(float) etc. in the reference CPU build.  Pinned
 *                     bit-exactly against the reference compiled from source.
 *   liboracle.so      transcendentals are the same IEEE-float-only routines as
 *                     the HIP kernels' (range reduction + polynomials), so HIP
 *                     and oracle trajectories agree bit for bit.
 *
 * Pinning summary (what checks each part; tests/ hold the checks):
 *   tables, triangle test, octree build + propagate, material preprocessing,
 *   the four physics plugins, boundary crossing and the whole CPU driver loop:
 *   bit-exact against vectors written by the reference compiled from source
 *   (oracle/ref/), plus closed forms (Moller-Trumbore on hand-made rays,
 *   brute-force triangle list for propagate, linear interpolation identities).
 *   mt19937: standard known answer (10000th output of seed 5489).
 *   Philox4x32-10: Random123 published known-answer vectors.
 */
#ifndef NBL_ORACLE_H
#define NBL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types -- */

typedef struct { float x, y, z; } orc_vec3;

/* 44-byte triangle, same layout as the reference's `triangle`
 * (core/triangle.h:70-77): materials first, then r0, e1=r1-r0, e2=r2-r0. */
typedef struct {
	int32_t mat_in;   /* side the normal points away from */
	int32_t mat_out;  /* side the normal points towards */
	orc_vec3 r0, e1, e2;
} orc_triangle;

/* 28-byte particle (core/particle.h:5-10). */
typedef struct { orc_vec3 pos, dir; float E; } orc_particle;

/* Interpolated lookup tables (common/util/table_{1,2,3}D.inl).  Row-major
 * [x][y][z], no pitch. */
typedef struct { const float* data; int n; float x_min, x_step; } orc_table1d;
typedef struct { const float* data; int nx, ny; float x_min, x_step, y_min, y_step; } orc_table2d;
typedef struct { const float* data; int nx, ny, nz;
	float x_min, x_step, y_min, y_step, z_min, z_step; } orc_table3d;

/* Special material indices (core/material_manager.h:77-86). */
enum {
	ORC_NOP = -128, ORC_TERMINATOR = -127, ORC_DETECTOR = -126,
	ORC_DETECTOR_LT50 = -125, ORC_DETECTOR_GE50 = -124, ORC_VACUUM = -123,
	ORC_MIRROR = -122
};

/* Inelastic model ids. */
enum { ORC_INEL_PENN = 0, ORC_INEL_KIEFT = 1 };

/* A material after the reference's create() preprocessing. */
typedef struct {
	float barrier;                 /* core/material.h:39-45 */
	/* kieft_elastic (physics/kieft/elastic.h) */
	orc_table1d el_imfp;           /* log(imfp nm^-1) vs log(E/eV) */
	orc_table2d el_icdf;           /* cos(theta) vs (log E, P) */
	float phonon_loss, recoil_const;
	/* inelastic: full_penn (physics/full_penn.h) or kieft_inelastic
	 * (physics/kieft/inelastic.h) */
	int inel_model;
	orc_table1d in_imfp;           /* log(imfp) */
	orc_table2d in_icdf;           /* log(omega) (penn) or log(omega') (kieft) */
	orc_table3d pn_q;              /* penn: q vs (log E, omega/E, P) */
	orc_table3d pn_ion;            /* penn: binding vs (log E, log(omega/E), P) */
	orc_table2d ki_ion;            /* kieft: binding vs (log omega', P) */
	float fermi, band_gap;
	/* Model parameters of the derivative estimator (no reference
	 * counterpart): sampled elastic / inelastic free paths are multiplied by
	 * these mean-free-path scales (1 after create). */
	float el_scale, in_scale;
} orc_material;

/* Physics options: template flags of physics_config.h:21-38 made runtime. */
typedef struct {
	int el_phonon_loss, el_atomic_recoil;           /* kieft_elastic<.,T,T> */
	int pn_generate_secondary;                      /* full_penn<.,T> */
	int ki_optical_phonon, ki_generate_secondary,   /* kieft_inelastic<...> */
	    ki_instantaneous_momentum, ki_momentum_conservation;
	int bi_quantum_transmission, bi_interface_refraction,
	    bi_interface_absorption;                    /* boundary_intersect<T,T,F> */
} orc_physics_opts;

/* Linearised octree geometry (geometry/octree.h:112-116). */
typedef struct {
	const int32_t* octree_data;
	const orc_triangle* triangles;
	orc_vec3 center, halfsize;
	float max_extent;
	int32_t n_octree, n_triangles;
} orc_geometry;

typedef struct {
	orc_geometry geom;
	const orc_material* materials;
	int n_materials;
	orc_physics_opts opts;
	float energy_threshold;
} orc_world;

/* Uniform random source.  kind 0: std::mt19937 + generate_canonical<float,24>
 * (common/util/random.inl:7-28).  kind 1: counter-based Philox4x32-10 stream
 * keyed by (seed) with counter (sid_lo, sid_hi, step, block), uniforms in
 * (0,1].  kind 2: a caller-supplied array of uniforms. */
typedef struct {
	int kind;
	uint32_t mt[624]; int mti;
	uint32_t key[2]; uint32_t ctr[4]; uint32_t buf[4]; uint32_t idx;
	const float* arr; size_t arr_n, arr_pos;
	uint64_t draws;
} orc_rng;

/* Per-particle state of the tracking loop. */
typedef struct {
	orc_particle p;
	int32_t mat;        /* current material index */
	int32_t last_tri;   /* index of last intersected triangle, -1 for none */
	uint32_t tag;       /* primary tag */
	uint32_t step;      /* step counter (Philox counter word 2) */
	uint64_t sid;       /* stream id (Philox counter words 0-1) */
} orc_state;

/* Status after a step. */
enum { ORC_ALIVE = 0, ORC_TERMINATED = 1, ORC_DETECTED = 2 };
/* Event chosen by init (status values of drivers/cpu/cpu_particle_manager.h). */
enum { ORC_EV_NONE = 0, ORC_EV_INELASTIC = 1, ORC_EV_ELASTIC = 2,
	ORC_EV_INTERSECT = 3 };

/* ------------------------------------------------------------- utilities -- */

int orc_math_mode(void);  /* 0 = C library float math, 2 = portable float math */
/* Evaluate one transcendental of this build: fn 0 log, 1 exp, 2 expm1,
 * 3 cbrt, 4 sin, 5 cos, 6 atan2(a, b). */
float orc_pf_eval(int fn, float a, float b);
void orc_pf_eval_batch(int fn, const float* a, const float* b, float* out, size_t n);

void orc_mt_seed(orc_rng* r, uint32_t seed);
uint32_t orc_mt_next(orc_rng* r);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_rng_init_mt(orc_rng* r, uint32_t seed);
void orc_rng_init_philox(orc_rng* r, uint64_t seed);
void orc_rng_set_stream(orc_rng* r, uint64_t sid, uint32_t step);
void orc_rng_init_array(orc_rng* r, const float* u, size_t n);
float orc_rng_unit(orc_rng* r);
uint64_t orc_se_sid(uint64_t parent_sid, uint32_t parent_step);

/* ---------------------------------------------------------------- tables -- */
float orc_t1_get(const orc_table1d* t, float x);
float orc_t2_get(const orc_table2d* t, float x, float y);
float orc_t2_get_rounddown(const orc_table2d* t, float x, float y);
float orc_t2_get_nearest(const orc_table2d* t, float x, float y);
float orc_t3_get(const orc_table3d* t, float x, float y, float z);
float orc_t3_get_rounddown(const orc_table3d* t, float x, float y, float z);
void orc_t1_probe(const orc_table1d* t, const float* x, float* out, size_t n);
void orc_t2_probe(const orc_table2d* t, int mode, const float* x, const float* y, float* out, size_t n);
void orc_t3_probe(const orc_table3d* t, int mode, const float* x, const float* y, const float* z, float* out, size_t n);

/* -------------------------------------------------------------- geometry -- */
float orc_intersect_ray(const orc_triangle* t, orc_vec3 start, orc_vec3 dir);
int orc_in_domain(const orc_geometry* g, orc_vec3 pos);
/* Returns the hit triangle index (or -1) and writes the distance. */
int32_t orc_propagate(const orc_geometry* g, orc_vec3 start, orc_vec3 dir,
	float distance, int32_t ignore_tri, int32_t ignore_mat, float* out_dist,
	uint32_t* out_cells);
int32_t orc_trilist_propagate(const orc_triangle* tris, int32_t n, orc_vec3 start,
	orc_vec3 dir, float distance, int32_t ignore_tri, int32_t ignore_mat, float* out_dist);
void orc_geometry_queries(const orc_geometry* g, const float* pos, const float* dir,
	const float* dist, const int32_t* ignore_tri, const int32_t* ignore_mat,
	int32_t* out_tri, float* out_dist, size_t n);
void orc_tri_intersect_batch(const orc_triangle* tris, const int32_t* tri_idx,
	const float* pos, const float* dir, float* out_t, size_t n);

/* Octree builder (geometry/octree/octree_builder.cpp).  Allocates outputs with
 * malloc; free them with orc_free.  Returns 0 on success. */
int orc_octree_build(const orc_triangle* tris, int32_t ntri,
	int32_t** octree_data, int32_t* n_octree,
	orc_triangle** sorted_tris, int32_t* n_sorted,
	float center[3], float halfsize[3]);
void orc_free(void* p);
int orc_tribox_overlap(const double center[3], const double halfsize[3],
	const double verts[9]);
/* Triangle constructor semantics (core/triangle.inl:1-4). */
void orc_make_triangle(const float r0[3], const float r1[3], const float r2[3],
	int32_t mat_in, int32_t mat_out, orc_triangle* out);

/* ------------------------------------------------- material preprocessing -- */
/* Raw dataset as read from a .mat file (io/hdf5_file.cpp): float data, the
 * dataset's unit in internal units (eV, nm, ...), and per axis the dimension
 * scale's first and last stored value times its unit, plus the count. */
typedef struct {
	const float* data;
	double unit;
	int ndim;
	int dims[3];
	double ax_front[3], ax_back[3];
	int ax_present[3];
} orc_raw_dataset;

/* Raw material (attributes in internal units: energies in eV, mass in the
 * reference's gram unit 6.2415e15). */
typedef struct {
	double barrier, fermi, band_gap, phonon_loss, effective_A;
	int has_band_gap;
	int inel_model;
	orc_raw_dataset el_imfp, el_icdf;
	orc_raw_dataset in_imfp, in_icdf;      /* penn: omega_icdf; kieft: w0_icdf */
	orc_raw_dataset pn_q, pn_ion;          /* penn */
	orc_raw_dataset ki_ion;                /* kieft 2-D binding_icdf */
	const double* outer_shells; int n_outer_shells; double outer_shells_unit;
} orc_raw_material;

/* Preprocess a raw material.  Table storage is malloc'ed and owned by *out;
 * release with orc_material_free.  Returns 0 on success, <0 on the same
 * validation failures the reference throws on. */
int orc_material_create(const orc_raw_material* raw, orc_material* out);
void orc_material_free(orc_material* m);
float orc_material_max_energy(const orc_material* m);

/* ---------------------------------------------------------------- physics -- */
/* scatter_list::sample_path (core/scatter_list.h:101-107). */
void orc_sample_path(const orc_material* m, const orc_particle* p, orc_rng* rng,
	uint8_t* type, float* distance);
/* Event bodies; each reads/writes *p.  `se` receives a secondary electron and
 * the return value says whether one was made. */
void orc_elastic_execute(const orc_material* m, const orc_physics_opts* o,
	orc_particle* p, orc_rng* rng);
int orc_penn_execute(const orc_material* m, const orc_physics_opts* o,
	orc_particle* p, orc_rng* rng, orc_particle* se);
int orc_kieft_execute(const orc_material* m, const orc_physics_opts* o,
	orc_particle* p, orc_rng* rng, orc_particle* se);
/* boundary_intersect::execute; returns ORC_ALIVE/TERMINATED/DETECTED and may
 * change *mat. */
int orc_boundary_execute(const orc_world* w, const orc_triangle* tri,
	orc_particle* p, int32_t* mat, orc_rng* rng);

/* One full step (init + its event).  semantics 0: CPU driver
 * (drivers/cpu/cpu_driver.inl: stored direction left unnormalised);
 * semantics 1: GPU driver (drivers/gpu/gpu_driver.inl:403-566: the
 * normalised direction is stored by init).  Returns ORC_ALIVE/TERMINATED/
 * DETECTED; *se_made says whether *se was filled.  *event gets the event. */
int orc_step(const orc_world* w, int semantics, orc_state* s, orc_rng* rng,
	orc_state* se, int* se_made, int* event);

/* ------------------------------------------- likelihood-ratio scores -- */
/* Parameters whose derivative is estimated: per material, the elastic and
 * inelastic mean-free-path scales and the barrier (work function + Fermi
 * energy).  pidx[m][k] is the parameter index of (material m, kind k) or -1.
 * A particle's score is the sum over its own steps, and over its ancestors'
 * steps up to its creation, of d log p(step outcome) / d theta:
 *   free path in material m, rates r_j = 1/tau_j (tau_j = scaled mean free
 *   path), travelled distance d, event j or none (boundary first):
 *       d/d s_j = (d * r_j - [event == j]) / s_j
 *   boundary transmission with T = 4s/(1+s)^2, s = sqrt(1 + dU/(E c^2)),
 *   dU = b_out - b_in:
 *       transmitted: d log T / d dU = (1-s) / (2 s^2 (1+s) E c^2)
 *       reflected:   d log(1-T) / d dU = 2 / (s dU)
 *   interface absorption (if enabled), P = exp(1 + E/(2 dU)):
 *       absorbed:    d log P / d dU = -E / (2 dU^2)
 *       not:         P E / (2 dU^2 (1-P))
 *   and d dU / d b_out = 1, d dU / d b_in = -1.
 * The kinematic effects of a barrier (energy shift, refraction) are not
 * random choices and carry no score. */
#define ORC_MAX_PARAMS 8
#define ORC_MAX_MATERIALS 16
enum { ORC_P_EL_SCALE = 0, ORC_P_IN_SCALE = 1, ORC_P_BARRIER = 2 };
typedef struct {
	int n_params;
	int pidx[ORC_MAX_MATERIALS][3];
} orc_score_map;

/* orc_step plus score accumulation into score[n_params] (map may be NULL:
 * then it is orc_step). */
int orc_step_scored(const orc_world* w, int semantics, orc_state* s, orc_rng* rng,
	orc_state* se, int* se_made, int* event, const orc_score_map* map, double* score);

/* ----------------------------------------------------------------- driver -- */
typedef struct {
	orc_particle p;
	uint32_t tag;
	uint64_t sid;
} orc_detected;

/* Reference CPU driver loop (drivers/cpu/simple_cpu_driver.h:53-65 with
 * cpu_mt_main.cpp:176-197), one thread, one mt19937 stream seeded with `seed`,
 * primaries pulled one at a time.  Detected records are written in the
 * reference's flush order.  Returns the number of records, or -1 if
 * `cap` was too small.  *n_steps receives the number of init calls. */
int64_t orc_run_reference(const orc_world* w, const orc_particle* primaries,
	size_t n, uint32_t seed, orc_detected* out, size_t cap, uint64_t* n_steps);

/* Philox-stream driver: every particle owns the stream (seed, sid, step), so
 * the result does not depend on scheduling.  Primary i gets sid = sid_base+i
 * and tag = tag_base+i.  Runs `threads` POSIX threads.  GPU semantics. */
int64_t orc_run_philox(const orc_world* w, const orc_particle* primaries,
	size_t n, uint64_t seed, uint64_t sid_base, uint32_t tag_base, int threads,
	orc_detected* out, size_t cap, uint64_t* n_steps);

/* Energy-deposit record (drivers/cpu/energydeposit_cpu_driver.h:55-66):
 * position after a scattering event that changed the energy, energy
 * before, energy lost, primary tag, stream id. */
typedef struct {
	orc_vec3 pos;
	float energy, deposit;
	uint32_t tag;
	uint64_t sid;
} orc_deposit;

/* orc_run_reference that also writes energy-deposit records in the order of
 * energydeposit_cpu_driver::simulate_to_end (energydeposit_cpu_driver.h:
 * 75-91); *n_dep receives their number (-1 is returned if above dep_cap). */
int64_t orc_run_reference_deposits(const orc_world* w, const orc_particle* primaries,
	size_t n, uint32_t seed, orc_detected* out, size_t cap, orc_deposit* dep, size_t dep_cap,
	uint64_t* n_dep, uint64_t* n_steps);

/* orc_run_philox that also writes the energy-deposit records (elastic and
 * inelastic events whose energy change is nonzero) into dep[dep_cap];
 * *n_dep receives their number (may exceed dep_cap: then -1 is returned). */
int64_t orc_run_philox_deposits(const orc_world* w, const orc_particle* primaries,
	size_t n, uint64_t seed, uint64_t sid_base, uint32_t tag_base, int threads,
	orc_detected* out, size_t cap, orc_deposit* dep, size_t dep_cap, uint64_t* n_dep,
	uint64_t* n_steps);

/* orc_run_philox with scores: out_scores[i * map->n_params + k] is the
 * score of detected record i for parameter k (secondaries inherit their
 * parent's score at creation). */
int64_t orc_run_philox_scored(const orc_world* w, const orc_particle* primaries,
	size_t n, uint64_t seed, uint64_t sid_base, uint32_t tag_base, int threads,
	const orc_score_map* map, orc_detected* out, double* out_scores, size_t cap,
	uint64_t* n_steps);

#ifdef __cplusplus
}
#endif
#endif /* NBL_ORACLE_H */
