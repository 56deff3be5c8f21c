/*
 * libfluid_amd — C ABI of the MI355X (gfx950) Position Based Fluids step and
 * its marching-cubes surface extraction.
 *
 * Plain C: opaque context, POD parameter blocks, plain pointers and sizes, an
 * int status on every call (0 = ok, < 0 = error; lf_last_error() has the
 * text). No exceptions cross this boundary.
 *
 * Each entry point names the reference interface it replaces
 * (tom91136/libfluid, paths relative to the reference root).
 *
 * Host pointers are host memory. Everything the step works on lives in device
 * memory owned by the context; it persists across calls and grows on demand.
 * Calls on one context are ordered on the context's own HIP stream; calls that
 * hand data back to the host synchronise it first. (The density-field
 * lf_mc_sample, lf_mc_classify_count and lf_mc_emit run on a second stream of
 * the context, created on first use and ordered by events after everything
 * issued before them, so they overlap the next step, which writes the other
 * of two alternating p* / cell-table buffers; every call that reads the mesh
 * or rewrites the field waits for them; lf_sync waits for both.)
 */
#ifndef LIBFLUID_AMD_H
#define LIBFLUID_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LF_OK 0
#define LF_ERR_ARG (-1)        /* bad argument (null pointer, size, ...) */
#define LF_ERR_HIP (-2)        /* HIP runtime error (text in lf_last_error) */
#define LF_ERR_MORTON (-3)     /* grid extent >= 1024 cells on an axis: the
                                  10-bit Morton axis of include/fluid/curves.h:56-72
                                  would wrap */
#define LF_ERR_CAPACITY (-4)   /* marching-cubes output exceeded its capacity */
#define LF_ERR_STATE (-5)      /* call out of order (e.g. lambda before grid) */
#define LF_ERR_NODEVICE (-6)   /* no HIP device */
#define LF_ERR_DIVERGED (-7)   /* the state diverged: a particle's p* is not
                                  finite, or a cell of the grid holds more
                                  particles than the guard allows (lf_set_guard);
                                  the counts are in lf_last_error and lf_health */

/* ---- particle types (include/fluid/fluid.hpp:18-20) --------------------- */
#define LF_FLUID 0
#define LF_OBSTACLE 1

/* ---- flags of lf_pbf_params.flags ---------------------------------------- */
#define LF_FLAG_DIFFUSE 0x1u   /* run the colour diffusion pass
                                  (sph_diffuse, oclsph_kernel.h:281-303) */
#define LF_FLAG_KAHAN 0x2u     /* Kahan-compensated density and lambda-denominator
                                  sums in the lambda pass (mixed precision) */
#define LF_FLAG_FAST_MATH 0x4u /* hardware sqrt and reciprocal (1 ulp) in the
                                  lambda and delta pair terms instead of the
                                  correctly rounded ones: results are no longer
                                  bit-identical to the CPU restatement, only
                                  close (mixed precision) */
#define LF_FLAG_PACKED16 0x8u  /* the lambda and delta passes gather neighbour
                                  positions as cell-relative 16-bit fixed-point
                                  offsets (unit h/65536, with the low 5 bits of
                                  the cell coordinate: 8 bytes instead of 16);
                                  relative positions are exact to one unit per
                                  axis: close to the FP32 path, not identical
                                  (mixed precision) */

/*
 * Parameters of one PBF step. Replaces ClSphConfig (oclsph_type.h:6-13), the
 * compile-time -DSPH_H (oclsph.hpp:170-178) and the constants hard-coded in
 * oclsph_kernel.h:20-31; h is a runtime value here.
 */
typedef struct lf_pbf_params {
  float h;              /* kernel radius = grid cell size (SPH_H) */
  float dt;             /* Config::dt (fluid.hpp:175) */
  float scale;          /* Config::scale: world = solver units * scale */
  uint32_t iterations;  /* Config::iteration: lambda/delta passes per step */
  float gravity[3];     /* Config::constantForce (an acceleration: F = m*g) */
  float min_bound[3];   /* Config::minBound, world units */
  float max_bound[3];   /* Config::maxBound, world units */
  float rest_density;   /* RHO = 6378 (oclsph_kernel.h:21) */
  float cfm_epsilon;    /* CFM_EPSILON = 600 (oclsph_kernel.h:25) */
  float corr_k;         /* CorrK = 1e-4 (oclsph_kernel.h:30) */
  float corr_dq;        /* CorrDeltaQ / h = 0.3 (oclsph_kernel.h:26) */
  float vd;             /* VD = 0.49, velocity damping (oclsph_kernel.h:20) */
  float xsph_c;         /* XSPH viscosity c: v += (c/rho0) sum m_j (v_j - v) W6;
                           0 = off (the reference has none) */
  uint32_t flags;       /* LF_FLAG_* */
  float density_tol;    /* > 0: density-error convergence check. After each
                           lambda pass the mean |rho/rho0 - 1| over the
                           scene's fluid particles is reduced in FP64 on the
                           device; once it is below density_tol the scene's
                           remaining lambda/delta passes of the step are
                           skipped. 0 = always run `iterations` passes (the
                           reference) */
} lf_pbf_params;

/* Fills p with the reference's constants for kernel radius h (dt 1/120,
 * scale 1, 4 iterations, no gravity, unit box, XSPH off). */
void lf_pbf_params_default(lf_pbf_params* p, float h);

/* A collider triangle in world units: geometry::Triangle<float>
 * (geometry.hpp) as MeshCollider holds it (fluid.hpp:121-126). */
typedef struct lf_triangle {
  float v0[3], v1[3], v2[3];
} lf_triangle;

/* A gravity well (fluid.hpp:137-144); applied in the predict pass exactly as
 * oclsph.hpp:306-317 (radius 75 world units, force clamped to +-10). */
typedef struct lf_well {
  float centre[3];
  float force;
} lf_well;

/* A particle source (fluid.hpp:146-159): every step adds a floor(sqrt(rate))
 * x ceil(sqrt(rate)) lattice of fluid particles, spacing h*scale/2, centred
 * on `centre` in x and z, mass 1, with the given velocity and colour
 * (oclsph.hpp:678-691). */
typedef struct lf_source {
  float centre[3];
  float velocity[3];
  uint32_t rate;
  uint32_t colour;
} lf_source;

/* A drain (fluid.hpp:161-169): every step erases the fluid particles whose
 * distance to `centre` is below `width` (oclsph.hpp:693-706; depth unused, as
 * in the reference). */
typedef struct lf_drain {
  float centre[3];
  float width;
  float depth;
} lf_drain;

/* Marching-cubes parameters. Replaces ClMcConfig (oclsph_type.h:39-44). */
#define LF_MC_FIELD_REFERENCE 0 /* mc_lattice metaball field, lattice from the
                                   grid extent (oclsph_kernel.h:421-530,
                                   oclsph.hpp:817-819) */
#define LF_MC_FIELD_DENSITY 1   /* SPH density field rho(x) = sum m W6(|x-x_j|)
                                   on an explicit S^3 lattice spanning the box */
typedef struct lf_mc_params {
  uint32_t field;            /* LF_MC_FIELD_* */
  float isolevel;            /* Config::isolevel */
  float resolution;          /* Config::resolution (reference field only) */
  float particle_size;       /* ClMcConfig::particleSize = 25 (oclsph.hpp:672) */
  float particle_influence;  /* ClMcConfig::particleInfluence = 0.5 (:673) */
  uint32_t samples[3];       /* lattice samples per axis (density field only) */
  uint32_t attributes;       /* 1 = also emit normals + colours */
} lf_mc_params;

/* One output triangle: the memory layout of geometry::MeshTriangle<float>
 * (geometry.hpp:41-44): 6 x vec3 then 3 packed ARGB colours, 84 bytes. */
typedef struct lf_mesh_triangle {
  float v0[3], v1[3], v2[3];
  float n0[3], n1[3], n2[3];
  uint32_t c[3];
} lf_mesh_triangle;

/* One particle: the memory layout of fluid::Particle<size_t,float>
 * (fluid.hpp:53-60), 48 bytes. */
typedef struct lf_particle {
  uint64_t id;
  int32_t type;
  float mass;
  float position[3];
  float velocity[3];
  uint32_t colour;
} lf_particle;

/* ---- context ------------------------------------------------------------ */
typedef struct lf_ctx lf_ctx;

/* Replaces the ocl::SphSolver constructor's context/program/queue set-up
 * (oclsph.hpp:170-189) and the check_size layout probe (oclsph.hpp:217-243,
 * oclsph_kernel.h:276-278): struct layouts are fixed at compile time. */
int lf_ctx_create(int device, lf_ctx** out);
int lf_ctx_destroy(lf_ctx* ctx);
const char* lf_last_error(const lf_ctx* ctx);
int lf_sync(lf_ctx* ctx);

/* ---- particle state (device resident) ----------------------------------- */
/* Upload n particles (SoA, vec3 packed as 3 floats). Replaces the SoA pack and
 * upload of oclsph.hpp:822-853. The upload order is the "caller order".
 * n < 2^28 per context (kernels address particles with 32-bit byte offsets;
 * 2^28 particles need about 240 GB of device memory). */
int lf_state_upload(lf_ctx* ctx, size_t n, const float* position,
                    const float* velocity, const float* mass,
                    const int32_t* type, const uint32_t* colour);
/* Download the current state in its current device order (after a step: the
 * Morton order of that step, as oclsph.hpp:874-892 leaves xs). Any pointer
 * may be NULL to skip that array. order[i] = upload index of particle i. */
int lf_state_download(lf_ctx* ctx, size_t n, float* position, float* velocity,
                      float* mass, int32_t* type, uint32_t* colour,
                      uint32_t* order);
int lf_state_size(const lf_ctx* ctx, size_t* n);
int lf_set_wells(lf_ctx* ctx, size_t n, const lf_well* wells);
/* Mesh colliders for the delta pass (collideTriangle2,
 * oclsph_collision.h:25-58, called from sph_delta, oclsph_kernel.h:364-368):
 * the triangles of all MeshColliders concatenated in order, as advance()
 * builds them (oclsph.hpp:796-807). Every delta pass tests each particle's
 * new position against every triangle, in order; at the first triangle whose
 * plane is within 40 world units and whose plane projection lies inside it,
 * the particle moves to the projection and its velocity is reflected about
 * the plane. n = 0 removes them. Every scene of a batch sees the same mesh. */
int lf_set_colliders(lf_ctx* ctx, size_t n, const lf_triangle* tris);
/* Device-resident sources, drains and spatial queries for lf_pbf_step (the
 * host pre-pass of advance(), oclsph.hpp:678-706 and :770-791, kept on the
 * device). Each step first appends the sources' particles at the end of the
 * state (upload indices continue after the largest one so far), then erases
 * the drained fluid particles keeping the order of the rest (one 4-byte
 * readback of the new count), then steps. Queries are answered right after
 * the step's grid build, from the colours the step started with, as the
 * reference does: per query point (world units) the fluid particles of its
 * own cell, their count and mean ARGB colour; lf_query_results copies the
 * last step's answers out (count 0 = no answer). Not for batches; a context
 * with sources or drains cannot capture a graph. n = 0 removes them. */
int lf_set_sources(lf_ctx* ctx, size_t n, const lf_source* sources);
int lf_set_drains(lf_ctx* ctx, size_t n, const lf_drain* drains);
int lf_set_queries(lf_ctx* ctx, size_t n, const float* points);
int lf_query_results(lf_ctx* ctx, size_t n, uint32_t* counts, uint32_t* colours);

/* ---- the memory-mapped file wire format ----------------------------------- */
/* Blocks of the packed little-endian records libfluid's sample application
 * writes to particles.mmf and triangles.mmf for its renderer
 * (samples/structures.hpp:391-425; records per structures.hpp:73-85 and
 * 234-258; include/fluid/wire.hpp restates the whole format): a 24-byte
 * header {int64 timestamp, u64 entries, u64 written}, then the records. The
 * records are packed on the device and copied into `dst` (host memory, e.g. a
 * shared mapping of the file); the header is stored with written = 0 before
 * the copy and with written = entries after it. *bytes receives the block
 * size; dst = NULL only asks for it; a smaller capacity is LF_ERR_CAPACITY
 * and writes nothing.
 * lf_wire_particles: the state in its current (sorted) order; id = the
 * particle's upload index (the order array of lf_state_download).
 * lf_wire_triangles: the last lf_mc_emit output, normals 0 and colours
 * 0xFFFFFFFF when it had no attributes (as lf_mc_download). */
#define LF_WIRE_HEADER_BYTES 24
#define LF_WIRE_PARTICLE_BYTES 44
#define LF_WIRE_MESH_TRIANGLE_BYTES 84
int lf_wire_particles(lf_ctx* ctx, void* dst, size_t capacity, int64_t timestamp, size_t* bytes);
int lf_wire_triangles(lf_ctx* ctx, void* dst, size_t capacity, int64_t timestamp, size_t* bytes);

/* ---- the PBF step, stage by stage --------------------------------------- */
/* Predict (advectAndCopy, oclsph.hpp:289-325), Morton key
 * (computeBoundAndZindex, oclsph.hpp:327-387), stable radix sort (replaces the
 * std::sort of oclsph.hpp:733-743), payload reorder and the dense cell-start
 * table (oclsph.hpp:746-767). */
int lf_grid_build(lf_ctx* ctx, const lf_pbf_params* p);
/* sph_diffuse (oclsph_kernel.h:281-303). */
int lf_pbf_diffuse(lf_ctx* ctx, const lf_pbf_params* p);
/* sph_lambda (oclsph_kernel.h:305-332). */
int lf_pbf_lambda(lf_ctx* ctx, const lf_pbf_params* p);
/* sph_delta (oclsph_kernel.h:335-384), Jacobi: reads p* and writes a second
 * buffer, instead of the reference's in-place update. */
int lf_pbf_delta(lf_ctx* ctx, const lf_pbf_params* p);
/* sph_finalise (oclsph_kernel.h:404-416). */
int lf_pbf_finalise(lf_ctx* ctx, const lf_pbf_params* p);
/* XSPH viscosity (absent in the reference; c = p->xsph_c, no-op when 0). */
int lf_pbf_xsph(lf_ctx* ctx, const lf_pbf_params* p);
/* nsteps whole steps: grid, [diffuse], iterations x (lambda, delta),
 * finalise, xsph (runSphKernel, oclsph.hpp:563-641, plus the host pre-pass). */
int lf_pbf_step(lf_ctx* ctx, const lf_pbf_params* p, uint32_t nsteps);

/* ---- marching cubes ------------------------------------------------------ */
/* Field sampling: mc_lattice (oclsph_kernel.h:421-530) for the reference
 * field, or the SPH density field. Needs lf_grid_build for the reference
 * field (cell table) and a finished step (final positions). */
int lf_mc_sample(lf_ctx* ctx, const lf_pbf_params* p, const lf_mc_params* m);
/* Case index + triangle count per cube and the device-side exclusive scan
 * of the counts (mc_size, oclsph_kernel.h:545-596, plus the host sum of
 * oclsph.hpp:461-465, done on the device). */
int lf_mc_classify_count(lf_ctx* ctx, const lf_mc_params* m);
/* Triangle emission at scanned offsets, in cube-index order (mc_eval,
 * oclsph_kernel.h:632-735; deterministic instead of atomic_inc). World
 * positions use p->scale; p may be NULL to keep the current scale. */
int lf_mc_emit(lf_ctx* ctx, const lf_pbf_params* p, const lf_mc_params* m);
/* Upload a scalar field on an S[0] x S[1] x S[2] lattice (x-major, sample
 * (i,j,k) at (origin + (i,j,k) * step) * scale world units) for
 * classify / emit / download: the GPU counterpart of the CPU mesher
 * surface::parameterise (surface.hpp:269-384). */
int lf_mc_field_upload(lf_ctx* ctx, const uint32_t samples[3], const float origin[3],
                       const float step[3], float scale, const float* field);
/* Triangle count of the last emit (synchronises). */
int lf_mc_count(lf_ctx* ctx, size_t* ntri);
/* Copy up to cap triangles out (synchronises). */
int lf_mc_download(lf_ctx* ctx, lf_mesh_triangle* out, size_t cap,
                   size_t* ntri);
/* Maximum triangles kept on the device (default 16M); LF_ERR_CAPACITY is
 * reported by lf_mc_count / lf_mc_download when an emit overflowed it. */
int lf_mc_set_capacity(lf_ctx* ctx, size_t max_triangles);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Marching cubes of a scalar field already in device
 * memory, with nothing allocated: d_field on a samples[0] x samples[1] x
 * samples[2] lattice (x-major, sample (i,j,k) at (origin + (i,j,k) * step) *
 * scale), the vertex positions of the first tri_cap triangles into d_tri_pos
 * (9 floats per triangle, cube-index order; the values lf_mc_field_upload +
 * lf_mc_classify_count + lf_mc_emit give), the full triangle count into
 * *d_ntri (device memory; above tri_cap means the output was cut). Enqueued
 * on the context's stream (lf_sync waits). The per-block counts live in a
 * fixed 1 MiB device array per GPU that every call on that GPU shares: calls
 * that overlap in time on one GPU (two contexts, two host threads) race. At
 * most 2^28 cubes ((S0-1)(S1-1)(S2-1)); larger lattices get LF_ERR_ARG. */
int lf_mc_mesh_static(lf_ctx* ctx, const uint32_t samples[3], const float origin[3],
                      const float step[3], float scale, float isolevel, const float* d_field,
                      float* d_tri_pos, size_t tri_cap, uint32_t* d_ntri);

/* ---- the drop-in step ----------------------------------------------------- */
/* The core of ocl::SphSolver::advance (oclsph.hpp:661-936) after sources and
 * drains: upload, one step, marching cubes on the reference field, readback.
 * xs is permuted into this step's Morton order and overwritten as in
 * oclsph.hpp:874-892. *ntri receives the triangle count; at most tri_cap
 * triangles are written to tris (tris may be NULL when tri_cap is 0). */
int lf_advance(lf_ctx* ctx, const lf_pbf_params* p, const lf_mc_params* m,
               lf_particle* xs, size_t n, lf_mesh_triangle* tris,
               size_t tri_cap, size_t* ntri, uint32_t* order);
/* (order may be NULL; else order[i] = input index of output particle i.)
 * When *ntri exceeds tri_cap, the triangles stay on the device and
 * lf_mc_download fetches them. */

/* ---- batches: many independent scenes per launch -------------------------
 * No reference counterpart: libfluid steps one scene per solver. Here the
 * scenes of a batch share every launch of a step. Scene s owns the particle
 * range [offsets[s], offsets[s+1]) (offsets = prefix sums of counts) in both
 * upload and sorted order, its own lf_pbf_params (bounds, scale, gravity,
 * xsph_c, ...) and its own slice of the cell table, so each scene's results
 * are those of stepping it alone. All scenes must share h, iterations and
 * flags (they select the launch sequence). The one difference from stepping a
 * scene alone is the position in the sorted order of particles whose cell
 * lies outside the scene's table (there are none while particles stay within
 * 2h of the bounds): they keep upload order instead of key order. */
int lf_batch_upload(lf_ctx* ctx, uint32_t n_scenes, const lf_pbf_params* params,
                    const size_t* counts, const float* pos, const float* vel,
                    const float* mass, const int32_t* type, const uint32_t* colour);
/* nsteps steps of every scene of the batch. */
int lf_batch_step(lf_ctx* ctx, uint32_t nsteps);
/* Scene count and offsets (n_scenes + 1 entries; offsets may be NULL). */
int lf_batch_info(const lf_ctx* ctx, uint32_t* n_scenes, size_t* offsets);
/* Scene of the following lf_mc_sample (params NULL) and lf_grid_info calls. */
int lf_batch_select(lf_ctx* ctx, uint32_t scene);

/* ---- FP64 diagnostics ------------------------------------------------------
 * Conserved quantities and the density error of one scene, reduced on the
 * device in FP64 (per-chunk partial sums, then a fixed-order sum per scene,
 * so the figures are reproducible). No reference counterpart. */
typedef struct lf_diag {
  uint64_t fluid;             /* fluid particles */
  double mass;                /* sum of fluid masses */
  double momentum[3];         /* sum of m v over fluid particles */
  double density_error_mean;  /* mean |rho/rho0 - 1| of the last lambda pass */
  double density_error_max;   /* max  |rho/rho0 - 1| of the last lambda pass */
  uint32_t passes;            /* lambda passes the scene ran in the last step */
  uint32_t pad;
  double pass_error[16];      /* with density_tol > 0: mean |rho/rho0 - 1| after
                                 each of the first 16 lambda passes of the last
                                 step (0 for passes not run) */
} lf_diag;
/* Diagnostics of scene `scene` (0 for a single-scene context) for the
 * current state (synchronises). */
int lf_diagnostics(lf_ctx* ctx, uint32_t scene, lf_diag* out);

/* ---- execution configuration ----------------------------------------------
 * threads: workgroup size of the particle kernels (128, 256 or 512; each
 * workgroup works on chunks of at most that many particles of one scene).
 * persistent: 1 = the pair kernels (list build, lambda, delta, XSPH) run a
 * device-sized grid that takes chunks from a work queue; 0 = one workgroup per
 * chunk. skin: neighbour-list skin as a fraction of h (0 turns the lists'
 * reuse off: every pass after the first enumerates cells). None of these
 * change results, only speed. */
int lf_exec_config(lf_ctx* ctx, uint32_t threads, int persistent, float skin);
/* The full execution configuration. particles_per_thread (1, 2 or 4): a
 * workgroup of a launch without the work queue takes that many consecutive
 * chunks, each thread one particle of each. mc_tile: the density field's
 * sample tile per workgroup of 256 threads, each thread a column of samples
 * along z: 0 = 8x8x16 samples (x, y, z; 4 per column), 1 = 4x16x16 (4 per
 * column), 2 = 8x8x8 (2 per column), 3 = 8x8x4 (1 per column). None of these
 * change results. */
typedef struct lf_exec_params {
  uint32_t threads;
  int32_t persistent;
  float skin;
  uint32_t particles_per_thread;
  uint32_t mc_tile;
} lf_exec_params;
int lf_exec_configure(lf_ctx* ctx, const lf_exec_params* e);
typedef struct lf_tune_result {
  uint32_t threads;
  int32_t persistent;
  float skin;
  float ms_per_step;   /* best configuration's mean step time */
  uint32_t trials;     /* configurations timed */
  uint32_t particles_per_thread;
  uint32_t mc_tile;
  float mc_ms;         /* best tile's density-field time (0: none sampled yet) */
} lf_tune_result;
/* Times `steps` steps (plus one warm-up) of the current state under each
 * configuration of threads x particles per thread (without the queue) or
 * persistent (with it) x skin, keeps the fastest, then, when the context has
 * sampled a density field before, times that field under each tile shape and
 * keeps the fastest. The state is restored after every trial: the call does
 * not advance the simulation. Refused while sources or drains are set. */
int lf_autotune(lf_ctx* ctx, uint32_t steps, lf_tune_result* out);

/* ---- HIP graphs --------------------------------------------------------------
 * Captures nsteps (even, so the double buffers return to their places) whole
 * steps, each followed by the marching cubes of m when m is not NULL (sample,
 * classify + count, emit), into one HIP graph; lf_graph_launch replays it
 * `times` times. Needs one eager step with the same sizes first (buffers are
 * not allocated during capture). Uploads and parameter changes drop the
 * graph. */
int lf_graph_capture(lf_ctx* ctx, uint32_t nsteps, const lf_mc_params* m);
int lf_graph_launch(lf_ctx* ctx, uint32_t times);

/* ---- inspection (parity tests, teacher forcing) ------------------------- */
#define LF_ARR_KEYS 1       /* u32[n]   sorted keys (scene table base + Morton key) */
#define LF_ARR_ORDER 2      /* u32[n]   upload index of each sorted particle */
#define LF_ARR_TABLE 3      /* u32[G]   dense cell-start table (all scenes' tables) */
#define LF_ARR_PSTAR 4      /* f32[n*4] p* (x,y,z,mass) in sorted order */
#define LF_ARR_LAMBDA 5     /* f32[n*4] (p*, lambda) of the last lambda pass */
#define LF_ARR_POSITION 6   /* f32[n*4] world position (x,y,z,mass) */
#define LF_ARR_VELOCITY 7   /* f32[n*4] velocity (x,y,z,colour bits) */
#define LF_ARR_TYPE 8       /* u8[n]    particle type */
#define LF_ARR_DIFFUSED 9   /* u32[n]   diffused ARGB colour */
#define LF_ARR_FIELD 10     /* f32[S]   MC field values (x-major lattice) */
#define LF_ARR_MCNORMAL 11  /* f32[S*4] MC field normal (x,y,z,0) */
#define LF_ARR_MCCOLOUR 12  /* u32[S]   MC field colour (ARGB bytes) */
#define LF_ARR_CASES 13     /* u8[M]    MC cube case index */
#define LF_ARR_DENSITY 14   /* f32[n]   density of the last lambda pass */
#define LF_ARR_LISTSTATS 15 /* u32[4]   neighbour lists: [0] float bits of the largest
                               displacement since the build, [1..3] particles that
                               enumerated cells instead (lambda, delta, XSPH) */
#define LF_ARR_CELLKEY 16   /* u32[n]   scene-local Morton cell key of each sorted particle */
int lf_read(lf_ctx* ctx, int array, void* dst, size_t bytes);
/* Keep the case index of every cube for lf_read(LF_ARR_CASES) (off by
 * default: 1 byte per cube of extra traffic). */
int lf_mc_keep_cases(lf_ctx* ctx, int on);
int lf_write(lf_ctx* ctx, int array, const void* src, size_t bytes);
/* Grid geometry of the last lf_grid_build: G, extent[3], min_extent[3]. */
int lf_grid_info(const lf_ctx* ctx, uint32_t* table_size, uint32_t extent[3],
                 float min_extent[3]);
/* MC lattice geometry of the last lf_mc_sample. */
int lf_mc_info(const lf_ctx* ctx, uint32_t samples[3], float origin[3],
               float step[3]);

/* ---- divergence guard -----------------------------------------------------
 * No reference counterpart: a diverged libfluid scene keeps stepping, ever
 * slower (a cell holding every particle makes each pass quadratic). Each
 * step's key pass counts the particles whose predicted p* is not finite and
 * the grid build counts the cells of the table holding more than
 * max_cell_particles particles (default 256; 0 = no limit). Once either count
 * is non-zero, the pair passes (list build, lambda, delta, finalise, XSPH
 * copies velocities through, diffusion) and the density field of later
 * launches do nothing, so the state stays as it was at the start of the step
 * in which it diverged (in that step's sorted order), and lf_pbf_step,
 * lf_batch_step and lf_graph_launch return LF_ERR_DIVERGED (with the counts in
 * lf_last_error) once the copy of the counts made after an earlier step call
 * has reached the host (lf_sync waits for it). The counts accumulate until the
 * next upload or lf_set_guard call. */
int lf_set_guard(lf_ctx* ctx, uint32_t max_cell_particles);
/* Synchronises; out[0] = particles with a non-finite p*, out[1] = cells above
 * the guard, out[2] = the largest such cell's particle count. */
int lf_health(lf_ctx* ctx, uint32_t out[3]);

/* ---- timing ----------------------------------------------------------------
 * With profiling on, every kernel launch of the context is bracketed by HIP
 * events on the context's stream; lf_kernel_stats reports per kernel name the
 * launch count and summed device time. */
int lf_profile_enable(lf_ctx* ctx, int on);
/* Times only the launches named `name` (as in lf_kernel_stats; NULL or "":
 * every launch). The two events around each launch cost ~8 % of a 1M step
 * when every launch is timed. */
int lf_profile_only(lf_ctx* ctx, const char* name);
int lf_profile_reset(lf_ctx* ctx);
/* Writes "name count total_ms\n" lines into buf (NUL terminated). */
int lf_kernel_stats(lf_ctx* ctx, char* buf, size_t cap);
/* Copy-kernel bandwidth probe (no reference counterpart: the denominator of
 * the roofline figures): copies `bytes` device to device with a plain
 * 16-byte-per-lane kernel ("copy_probe") reps times; *ms = mean per copy. */
int lf_copy_probe(lf_ctx* ctx, size_t bytes, int reps, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* LIBFLUID_AMD_H */
