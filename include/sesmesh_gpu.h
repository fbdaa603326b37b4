/* sesmesh_gpu.h: C ABI of the MI355X (gfx950) mesh stage of sesmesh.
 *
 * The drop-in boundary for sesmesh's `--backend hip` mesh stage. The
 * reference crosses it in Rust through the `sesmesh_gpu::batch` traits
 * (`DeviceResident`, `DeviceBatch`) implemented by `hip::Resident` /
 * `hip::Batch`; each entry point below cites the reference interface it
 * replaces (paths relative to the reference repository root).
 *
 * Conventions: every function returns 0 on success or a negative SMG_E_*
 * code, never throws and never aborts; `smg_last_error()` gives a
 * thread-local message for the last failure. Pointers in *_ref structs are
 * host pointers read during the call only. Output views stay valid until the
 * next call that writes the same object. Thread safety: one thread drives a
 * resident at a time; a device handle may be shared.
 */
#ifndef SESMESH_GPU_H
#define SESMESH_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes, mirroring crates/sesmesh-gpu/src/error.rs:10-66. */
#define SMG_OK 0
#define SMG_E_NO_RUNTIME (-1)  /* Error::NoRuntime: HIP runtime unusable */
#define SMG_E_NO_DEVICE (-2)   /* Error::NoDevice: no such device ordinal */
#define SMG_E_UNSUPPORTED (-3) /* Error::Unsupported: not gfx950, or a
                                  parameter the device path does not take */
#define SMG_E_CALL (-4)        /* Error::Call: a HIP call failed */
#define SMG_E_SELFTEST (-5)    /* Error::SelfTest: device arithmetic differs
                                  from IEEE on the host */
#define SMG_E_LENGTH (-6)      /* Error::Length: array sizes inconsistent */
#define SMG_E_INVALID (-7)     /* invalid input (sesmesh Error::InvalidInput) */
#define SMG_E_INTERNAL (-8)    /* internal inconsistency (Error::Internal) */
#define SMG_E_CANCELLED (-9)   /* the caller's cancellation flag was set (Error::Cancelled) */

typedef struct smg_device smg_device;
typedef struct smg_arrangement smg_arrangement;
typedef struct smg_resident smg_resident;
typedef struct smg_mesh smg_mesh;
typedef struct smg_plan smg_plan;
typedef struct smg_model smg_model;
typedef struct smg_batch smg_batch;

/* Thread-local message of the last failure (never NULL). */
const char* smg_last_error(void);

/* ------------------------------------------------------------------ device
 * Replaces hip::Runtime::load/open (crates/sesmesh-gpu/src/hip/mod.rs:92-234)
 * and its open-time self-test (hip/selftest.rs:21-70): opens device
 * `ordinal`, checks it is gfx950, and runs the f32 self-test of the
 * reference (selftest.rs:25-38) plus an f64 one (the f64 exact refinement
 * runs on this device); a mismatch returns SMG_E_SELFTEST. */
typedef struct {
    char name[256];
    char arch[64];
    int compute_units;
    int clock_khz;
    size_t total_memory;
    size_t lds_per_block;
    int wavefront;
} smg_device_info;

int smg_device_open(int ordinal, smg_device** out, smg_device_info* info);
void smg_device_close(smg_device* dev);

/* Cancellation and progress of the device's runs, the reference's Control
 * (crates/sesmesh/src/control.rs:13-131) passed to compute_surface_with.
 * With `cancel` set (non-zero), a run returns SMG_E_CANCELLED promptly:
 * the library polls the flag between passes and while it waits, and once
 * it is set, raises a flag on the device that stops every running pass
 * from taking more work. `progress` (may be NULL) is called from the
 * calling thread with the reference's stages: (0, total) when a stage
 * starts and (total, total) when it ends; the mesh stage's units are
 * surface bricks, reported after sampling and after refinement too.
 * Planned runs report when they end. NULL clears the control. */
#define SMG_STAGE_GEOMETRY 0
#define SMG_STAGE_CLASSIFY 1
#define SMG_STAGE_MESH 2
#define SMG_STAGE_ASSEMBLE 3
#define SMG_STAGE_MEASURE 4
typedef void (*smg_progress_fn)(void* user, int32_t stage, uint64_t done, uint64_t total);
typedef struct {
    const volatile int32_t* cancel;
    smg_progress_fn progress;
    void* user;
} smg_control;
int smg_device_set_control(smg_device* dev, const smg_control* control);

/* Self-test on `n` seeded input pairs (selftest.rs:64-80 input generator):
 * counts of f32 and f64 results whose bits differ from the host's. */
int smg_selftest(smg_device* dev, size_t n, uint64_t seed, uint64_t* f32_mismatch, uint64_t* f64_mismatch);

/* ------------------------------------------------------------ arrangement
 * Host-side SAS arrangement of a sphere set for one probe radius, f64, as
 * crates/sesmesh/src/geom/arrangement.rs:437-645 builds it (with circles and
 * triple points, without arcs) plus the field's circle and vertex cell grids
 * (crates/sesmesh/src/field/point.rs:213-274). `xyzr` holds n rows of centre
 * x, y, z and radius; radius 0 is ignored. `threads` <= 0 means all cores.
 * The caller of the reference builds this on the CPU before the mesh stage;
 * it is provided here so a user of this library has the whole path.
 * smg_arrangement_build is the SES kind; probe 0 gives the vdW kind
 * (SesField::new_with, point.rs:224-228). */
int smg_arrangement_build(const double* xyzr, size_t n, double probe, int threads, smg_arrangement** out);
void smg_arrangement_free(smg_arrangement* a);

/* Surface kinds (SurfaceKind, crates/sesmesh/src/field/mod.rs:33-41). The SAS
 * and vdW surfaces are unions of spheres: their arrangements carry no
 * circles or SAS vertices (build_with(.., with_circles = false, ..),
 * point.rs:229-237); the vdW one is built with probe 0, whatever `probe`
 * says (SurfaceParams::effective_probe, surface.rs:411-417). */
#define SMG_KIND_SES 0
#define SMG_KIND_SAS 1
#define SMG_KIND_VDW 2
/* SesField::new(spheres, probe, kind) (point.rs:208-274) on the host. */
int smg_arrangement_build_kind(const double* xyzr, size_t n, double probe, int kind, int threads,
                               smg_arrangement** out);

typedef struct {
    const double* origin; /* 3 */
    double cell;
    uint32_t dims[3];
    const uint32_t* starts; /* dims[0]*dims[1]*dims[2] + 1 */
    const uint32_t* items;
    size_t n_items;
} smg_grid_ref;

/* The arrangement as the mesh stage reads it. Replaces ArrangementRef /
 * GridRef (crates/sesmesh-gpu/src/batch.rs:44-82), which carried f32
 * frame-local copies; this boundary carries the f64 arrangement itself and
 * the device derives the f32 copies, so its decisions are the CPU's. */
typedef struct {
    size_t n_spheres;
    const double* center;   /* 3 per sphere, internal (Morton) order */
    const double* radius;   /* vdW radius */
    const double* expanded; /* radius + probe */
    const uint8_t* status;  /* 0 exposed, 1 buried, 2 contained */
    const uint32_t* orig;   /* internal -> input index */
    const uint32_t* nb_off; /* n_spheres + 1: neighbours (exposed rows) */
    const uint32_t* nb_items;
    size_t n_circles;
    const uint32_t* circle_ij; /* 2 per circle */
    const double* circle;      /* 10 per circle: centre, axis, e1, radius */
    const uint32_t* cb_off;    /* n_circles + 1: circle blockers */
    const uint32_t* cb_items;
    size_t n_vertices;
    const double* vertex;        /* 3 per SAS vertex */
    const uint32_t* vertex_atoms; /* 3 per SAS vertex */
    smg_grid_ref sphere_grid, circle_grid, vertex_grid;
    double probe, max_expanded, rho_max;
    int32_t kind; /* SMG_KIND_* */
} smg_arrangement_ref;

/* View of a built arrangement (valid while it lives). */
int smg_arrangement_view(const smg_arrangement* a, smg_arrangement_ref* out);

/* --------------------------------------------------------------- resident
 * Replaces hip::Resident::new (crates/sesmesh-gpu/src/hip/resident.rs:343-371)
 * and DeviceResident::set_refinement (batch.rs:165-207): uploads the
 * arrangement once into HBM (SoA, f64 plus f32 copies). */
int smg_resident_new(smg_device* dev, const smg_arrangement_ref* arr, smg_resident** out);
void smg_resident_free(smg_resident* r);
/* DeviceResident::resident_bytes (batch.rs:165-207). */
size_t smg_resident_bytes(const smg_resident* r);

/* SesField::new(spheres, probe, kind) and Resident::new in one call, the
 * arrangement built on the device (geom/arrangement.rs:437-645, with the
 * field's cell grids, point.rs:213-274): the same arrays, bit for bit and
 * in the same order, as smg_arrangement_build_kind followed by
 * smg_resident_new, without the host build. `xyzr` (n rows of centre x, y,
 * z and radius) is read during the call. The one comparison the device
 * cannot repeat to the last bit, an exposed-arc angle gap from atan2 within
 * 1e-13 rad of the test's threshold, sends that sphere set to the host
 * builder instead (smg_resident_info reports which ran). */
int smg_resident_build(smg_device* dev, const double* xyzr, size_t n, double probe, int kind, smg_resident** out);

typedef struct {
    size_t n_spheres, n_circles, n_vertices; /* after containment; circles and SAS vertices of the SES kind */
    int32_t kind;                            /* SMG_KIND_* */
    int32_t built_on_device;                 /* 1: smg_resident_build's device path; 0: a host arrangement */
    double probe;                            /* the arrangement's probe (0 for vdW) */
    double build_ms;                         /* wall time of smg_resident_build */
    size_t bytes;                            /* HBM held */
} smg_resident_info;
int smg_resident_info_get(const smg_resident* r, smg_resident_info* info);

/* The resident's arrangement copied back to the host, as a new
 * smg_arrangement (free with smg_arrangement_free; view with
 * smg_arrangement_view). */
int smg_resident_arrangement(const smg_resident* r, smg_arrangement** out);

/* ------------------------------------------------------------ mesh stage
 * Parameters of a mesh stage: SurfaceParams' spacing, brick size and
 * refinement mode (crates/sesmesh/src/surface.rs:195-300). */
#define SMG_REFINE_EXACT 0 /* f64 refinement, CPU-identical bits */
#define SMG_REFINE_FAST 1  /* f32 refinement (not identical) */
typedef struct {
    double spacing;
    uint32_t brick;       /* even, 4..=32 (2x2x2 sub-blocks) */
    int32_t refine;       /* SMG_REFINE_* */
    int32_t sizing;       /* 0: lists tried from the first LDS tier; 1: from the tier the device's sizing model predicts */
    int32_t reserved;
} smg_mesh_params;

/* Per-pass device times (ms) of the last mesh stage, from HIP events; the
 * reference's BatchTimes/RefineTimes (batch.rs:118-161). */
typedef struct {
    double classify_ms;  /* touched bricks + brick and sub-block classification */
    double sample_ms;    /* candidate gather + field sampling */
    double edges_ms;     /* crossing edges, vertex numbering */
    double refine_ms;    /* vertex refinement + attributes */
    double cells_ms;     /* cell pass + assembly */
    double total_ms;
    /* Durations of the main kernels alone (HIP events around their
       launches on the library's stream). */
    double k_classify_ms, k_substates_ms, k_sample_ms, k_refine_ms;
} smg_mesh_times;

/* The whole mesh stage on the device: the work of surface.rs:1047-1079
 * (classification, mesh_stage, assemble) for the CPU back end, whose result
 * it reproduces bit for bit in SMG_REFINE_EXACT mode: surface bricks,
 * samples, vertices (position, normal, atom, patch) in brick order and
 * triangles as global vertex indices. The mesh stays in HBM; copy it out
 * with smg_mesh_copy. */
int smg_mesh_new(smg_resident* r, const smg_mesh_params* p, smg_mesh** out, smg_mesh_times* times);
void smg_mesh_free(smg_mesh* m);

typedef struct {
    size_t touched_bricks, surface_bricks, vertices, triangles;
    uint64_t refine_evaluations;
    size_t sampled_subblocks, overflow_subblocks;
    uint64_t evaluated_samples; /* samples the SES lane kernel evaluated */
    int64_t frame[3];
    double band, edge_clamp;
    /* Of the whole run (every part holds the same): sub-blocks a sampling or
       refinement pass took after an earlier pass's lists could not hold them. */
    uint64_t sample_retries, refine_retries;
} smg_mesh_info;
int smg_mesh_info_get(const smg_mesh* m, smg_mesh_info* info);

/* Copies the mesh to host arrays (any may be NULL): positions and normals
 * 3 f32 per vertex, atom (input index) u32 and patch u8 (0 convex, 1 saddle,
 * 2 concave) per vertex, triangles 3 u32 each, surface brick keys 3 i32
 * each, brick samples ((b+1)^3 f32 per brick, x fastest), per-vertex edge
 * parameter t (f64, unclamped) and evaluations (| 1<<31 if Newton
 * converged; | 1<<30 in fast mode for an edge whose f32 search did not
 * converge and was redone by the exact f64 search, as the reference redoes
 * such device edges on the CPU, crates/sesmesh/src/mesh/extract.rs:336-358). */
int smg_mesh_copy(const smg_mesh* m, float* positions, float* normals, uint32_t* atoms, uint8_t* patches,
                  uint32_t* triangles, int32_t* bricks, float* values, double* t, uint32_t* evals);

/* The exact clamped field at n points (xyz f64 each): for the SES kind
 * values[i] = probe - ses_eval(x, probe + band).distance, the distance
 * being 0 outside the SAS; for the SAS and vdW kinds sphere_eval(x, band),
 * the least |x - c| - R below band; f64, bit for bit the CPU's
 * SesField::clamped_value (crates/sesmesh/src/field/point.rs:759-765,
 * ses_eval :384-411, sphere_eval :742-756). Used to measure how far a
 * fast-mode vertex lies from the exact surface. */
int smg_field_clamped(smg_resident* r, const double* points, size_t n, double band, double* values);

/* Batched mesh stage: meshes `count` resident molecules of one device in one
 * pass (every kernel runs once over the concatenated work of all of them;
 * params[i] for residents[i], one brick size for the run). Part i of the
 * result is byte for byte the mesh smg_mesh_new makes of residents[i]
 * alone: the same bricks, samples, vertices and triangles (vertex indices
 * local to the part). The reference has no multi-molecule call; this
 * replaces a loop of `fold_resident` runs (crates/sesmesh/src/field/
 * gpu.rs:532-712) over molecules, the caller of batched serving.
 * smg_mesh_new(r, p, ...) is smg_mesh_many(&r, p, 1, ...), and
 * smg_mesh_info_get / smg_mesh_copy read part 0. */
int smg_mesh_many(smg_resident* const* residents, const smg_mesh_params* params, size_t count, smg_mesh** out,
                  smg_mesh_times* times);
size_t smg_mesh_count(const smg_mesh* m);
int smg_mesh_part_info(const smg_mesh* m, size_t part, smg_mesh_info* info);
int smg_mesh_part_copy(const smg_mesh* m, size_t part, float* positions, float* normals, uint32_t* atoms,
                       uint8_t* patches, uint32_t* triangles, int32_t* bricks, float* values, double* t,
                       uint32_t* evals);

/* ---------------------------------------------------------- measure stage
 * The measure stage of every part of a mesh, on the device: validation,
 * connected components and triangles grouped by component, as the CPU back
 * end runs them after assembly (crates/sesmesh/src/surface.rs:1081-1095):
 * mesh::validate_with (mesh/mod.rs:632-705), mesh::components_with
 * (mod.rs:181-300) and post::group_by_component (mesh/post.rs:14-34). The
 * f64 areas and volumes are summed in the reference's fixed chunk order, so
 * every component's sums and the stats are the CPU's bit for bit.
 * Afterwards the mesh's triangles are grouped by component (smg_mesh_copy
 * returns them in that order, as SurfaceOutput::mesh holds them). A mesh
 * with an index out of range is validated only (measured = 0). Measuring
 * again returns at once. `ms` (may be NULL) receives the device time. */
typedef struct {
    uint64_t boundary_edges, non_manifold_edges, misoriented_edges, degenerate_triangles, duplicate_triangles,
        invalid_indices, bad_vertices;
} smg_validation; /* mesh::Validation (mod.rs:579-597) */
typedef struct {
    uint64_t vertices, triangles;
    double area, volume; /* A^2; signed A^3, negative for a cavity */
    int64_t euler;       /* V - E + F */
    uint64_t first_triangle;
} smg_component; /* mesh::Component (mod.rs:131-149) */
typedef struct {
    smg_validation validation;
    uint64_t components, cavities;
    double area, volume; /* SurfaceStats::area / volume: sums over the components */
    int64_t euler;
    int32_t measured; /* 1: components measured and triangles grouped */
    int32_t pad;
} smg_measure_info;
int smg_mesh_measure(smg_mesh* m, double* ms);
int smg_mesh_part_measure(const smg_mesh* m, size_t part, smg_measure_info* out);
/* The part's components (cap entries at most) and per-vertex component ids
 * (either may be NULL). */
int smg_mesh_part_components(const smg_mesh* m, size_t part, smg_component* comps, size_t cap,
                             uint32_t* vertex_component);
/* A host mesh (one part) on the device, to measure or write it: n vertices
 * (positions 3 f32 each; normals, atoms and patches may be NULL: no
 * normals, zero atoms and patches), t triangles (3 u32 each). */
int smg_mesh_upload(smg_device* dev, const float* positions, const float* normals, const uint32_t* atoms,
                    const uint8_t* patches, size_t n, const uint32_t* triangles, size_t t, smg_mesh** out);

/* ---------------------------------------------------------------- writers
 * Mesh files of one part, formatted on the device, byte for byte as the
 * reference's writers: write_obj (crates/sesmesh/src/export/obj.rs:
 * 138-224, numbers by fmt.rs:30-68), write_ply (export/ply.rs:97-182,
 * binary little-endian) and write_msms (export/msms.rs:70-148, the .vert
 * and .face files as two formats). A measured mesh (smg_mesh_measure) is
 * the one compute_surface returns: triangles grouped by component, and the
 * PLY `component` property and MSMS face numbers from its components;
 * unmeasured, PLY has no component property and MSMS numbers every vertex
 * and face 1, as for a reference Mesh without component ids. */
#define SMG_SESMESH_VERSION "1.3.2" /* the version line of the files: sesmesh's (Cargo.toml:23) */
#define SMG_FORMAT_OBJ 0
#define SMG_FORMAT_PLY 1
#define SMG_FORMAT_MSMS_VERT 2
#define SMG_FORMAT_MSMS_FACE 3
typedef struct {
    int32_t precision;        /* OBJ decimals 0..=9 (ObjOptions::precision, default 4) */
    int32_t normals;          /* OBJ vn records and f a//a, PLY normals (default 1; only if the mesh has them) */
    const char* object_name;  /* OBJ `o` record (NULL: "ses") */
    const char* const* comments; /* OBJ `# ` / PLY `comment ` lines after the version line */
    size_t n_comments;
    const float* vertex_colors;  /* OBJ / PLY colours, 3 per vertex of the part (NULL: none) */
    const uint64_t* group_first; /* OBJ `g` records: first triangle, ascending */
    const char* const* group_names;
    size_t n_groups;
    const char* description;  /* MSMS header (MsmsHeader) */
    uint64_t spheres;
    double density, probe_radius;
} smg_export_options;
void smg_export_defaults(smg_export_options* o);
/* Formats part `part` in `format` into buf (cap bytes); *len receives the
 * file's size. With buf NULL only the size is computed; a buffer smaller
 * than the size returns SMG_E_LENGTH. */
int smg_export(const smg_mesh* m, size_t part, int format, const smg_export_options* o, char* buf, size_t cap,
               size_t* len);
int smg_export_file(const smg_mesh* m, size_t part, int format, const smg_export_options* o, const char* path);

/* --------------------------------------------------------- no allocation
 * The mesh stage of one molecule with no allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * The caller passes device buffers for the outputs, with their capacities;
 * every intermediate buffer comes from a fixed scratch array of
 * SMG_NOALLOC_SCRATCH_BYTES that each device holds for the life of the
 * process (set up at smg_device_open). The call makes no hipMalloc or
 * hipFree. The mesh written is smg_mesh_new's, byte for byte. A molecule
 * whose run needs more scratch than the array holds, or more vertices or
 * triangles than the output capacities, returns SMG_E_LENGTH (the outputs
 * are then unspecified). Calls on one device share its array and are
 * serialised by the library; calls that overlap on one device from
 * outside it (another library copy or direct writes) race. */
#define SMG_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
typedef struct {
    float* positions;   /* device, 3 per vertex */
    float* normals;     /* device, 3 per vertex */
    uint32_t* atoms;    /* device, 1 per vertex */
    uint8_t* patches;   /* device, 1 per vertex */
    uint32_t* triangles; /* device, 3 per triangle */
    uint64_t vertex_capacity, triangle_capacity;
} smg_mesh_out;
/* info (may be NULL) receives the counts of the mesh written. */
int smg_mesh_into(smg_resident* r, const smg_mesh_params* p, const smg_mesh_out* out, smg_mesh_info* info);

/* ----------------------------------------------------------------- tuning
 * Launch shapes of the mesh stage on a device. Every setting gives the same
 * meshes, byte for byte; the settings trade LDS capacity, occupancy and
 * lanes per edge. No counterpart in the reference, whose launch shapes are
 * fixed (crates/sesmesh-gpu/src/hip/resident.rs:675-683). */
typedef struct {
    int32_t sample_first_tier; /* 0..3: LDS capacity tier the sampling passes start at */
    int32_t refine_first_tier; /* 0..3: the same for refinement */
    int32_t refine_group_max;  /* power of two <= 64: most lanes refining one edge */
    int32_t uniform_shared;    /* 1: sub-block states from one shared gather per brick; 0: a wavefront per brick */
    int32_t scratch_workgroups; /* workgroups of the global-scratch passes, 16..8192 (0: 1024); ~0.9 MB each */
} smg_tuning;
/* Sets the tuning of the device's next runs (NULL: the defaults). */
int smg_device_set_tuning(smg_device* dev, const smg_tuning* tuning);
int smg_device_get_tuning(const smg_device* dev, smg_tuning* tuning);
/* Autotuner: times the mesh stage of these molecules (device time, best of
 * `repeats` runs) under candidate tunings, one knob at a time from the
 * device's current tuning, keeps the fastest on the device and returns it
 * with its time in ms. Plans capture the tuning current when they record. */
int smg_tune(smg_resident* const* residents, const smg_mesh_params* params, size_t count, int repeats,
             smg_tuning* best, double* best_ms);

/* ---------------------------------------------------------- sizing model
 * Learned per-brick sizing. No counterpart in the reference, which sizes
 * its lists by counting.
 *
 * The model is an MLP with layers 64 -> 128 -> 128 -> 2 and ReLU between
 * them: 25 090 parameters. Layers 1 and 2 run on the BF16 matrix cores with
 * FP32 accumulation; the rest is FP32. It reads 64 features of a brick: the
 * square roots of the sphere counts in 4 radial shells x 16 directions
 * around the brick centre, within reach (brick half-diagonal + 2 probe +
 * largest expanded radius). It predicts two things: an occupancy logit (is
 * the brick a surface brick) and log(1 + candidate load). The candidate load
 * is max(S + C, 2 near, 4 SAS vertices, 4 inside) of the brick's largest
 * sampling lists.
 *
 * Runs with params.sizing = 1 send each sub-block first to the LDS tier
 * its predicted load calls for (smg_model_set_policy). The exact gathers check every list,
 * so a wrong prediction costs one retry at the next tier and the mesh is
 * unchanged. smg_mesh_info counts the retries.
 *
 * Training is on the device, with this library's own forward and backward
 * kernels and Adam. The samples are touched bricks, labelled by the exact
 * classifier (occupancy) and the exact gathers (load). */
typedef struct {
    uint64_t samples;
    uint32_t epochs, batch;
    double occupancy_loss;     /* mean binary cross-entropy */
    double occupancy_accuracy; /* fraction of bricks whose logit sign matches */
    double load_rmse;          /* RMS error of log(1 + load) over surface bricks */
    double train_ms;
} smg_train_stats;

int smg_model_new(smg_device* dev, uint64_t seed, smg_model** out); /* He-uniform weights from `seed` */
void smg_model_free(smg_model* m);
size_t smg_model_param_count(void); /* 25090: W1 [64][128], b1, W2 [128][128], b2, W3 [128][2], b3 */
int smg_model_get_params(const smg_model* m, float* params);
int smg_model_set_params(smg_model* m, const float* params); /* resets the optimiser state */
/* Runs the exact mesh stage of these molecules and appends their touched
 * bricks (features, occupancy, log(1 + load)) to the training set. */
int smg_model_collect(smg_model* m, smg_resident* const* residents, const smg_mesh_params* params, size_t count);
size_t smg_model_samples(const smg_model* m);
/* Copies up to `cap` samples of the training set (any array may be NULL). */
int smg_model_dataset(const smg_model* m, float* features, uint8_t* occupied, float* log_load, size_t cap, size_t* n);
/* Outputs (occupancy logit, log(1 + load)) of n feature rows. */
int smg_model_forward(smg_model* m, const float* features, size_t n, float* out);
/* Parameter gradients of the mean loss over n rows: per row, BCE(occupancy)
 * + 0.5 occupied (predicted - log_load)^2. loss[0], loss[1] receive the two
 * mean terms. No update. */
int smg_model_gradients(smg_model* m, const float* features, const uint8_t* occupied, const float* log_load,
                        size_t n, float* grads, double* loss);
/* Adam (beta 0.9, 0.999) over shuffled batches of the training set. */
int smg_model_train(smg_model* m, uint32_t epochs, uint32_t batch, float lr, uint64_t seed, smg_train_stats* stats);
/* Placement: a sub-block goes first to the smallest tier holding scale x
 * its predicted load, up to tier max_tier (0..4). Default scale 0.3, tier 1:
 * larger tiers run at lower occupancy, so only sub-blocks predicted well
 * past the first tier skip it. */
int smg_model_set_policy(smg_model* m, float scale, int32_t max_tier);
/* The model runs with sizing 1 on this device use (NULL: none). */
int smg_device_set_model(smg_device* dev, smg_model* m);

/* ------------------------------------------------------------------ plans
 * A capture-ready mesh stage for serving loops. No counterpart in the
 * reference: it is the loop a service runs around fold_resident
 * (crates/sesmesh/src/field/gpu.rs:532-712), one call per batch of
 * molecules.
 *
 * A plan fixes capacities for its runs. Its buffers are allocated once, on
 * the first run, and the whole pipeline is then captured into a hipGraph:
 * every count stays in device memory, and persistent workgroups take work
 * from device-side queues. Each later run uploads the run's molecule table
 * and replays the graph, with one synchronisation at the end.
 *
 * A run whose counts exceed the capacities is redone by smg_mesh_many, the
 * exact count-then-allocate path. Such counts are found on the host for
 * molecules, spheres and brick boxes, and on the device for surface bricks,
 * vertices and triangles. The plan then grows to hold that run, and its
 * buffers and graph are rebuilt on the next run. The mesh is the same
 * either way: byte for byte what smg_mesh_many makes.
 *
 * The mesh of a run that fits views the plan's buffers. It must be freed
 * (smg_mesh_free) before the plan runs again; if the plan is freed first,
 * the view is detached and its copies fail. A run that falls back returns
 * a mesh of its own. */
typedef struct {
    uint32_t molecules; /* per run (table rows) */
    uint32_t brick;     /* brick size of every run */
    int32_t refine;     /* SMG_REFINE_* of every run */
    uint32_t pad;
    uint64_t locals;         /* sum over molecules of max(spheres, circles, SAS vertices) */
    uint64_t spheres;        /* sum of spheres */
    uint64_t box_bricks;     /* sum of the molecules' brick-box cells */
    uint64_t surface_bricks; /* device counts: */
    uint64_t vertices;
    uint64_t triangles;
} smg_plan_caps;

typedef struct {
    uint64_t runs, replays, fallbacks, captures;
    uint32_t last_overflow; /* 1 surface bricks, 2 vertices, 4 triangles, 8 table */
    uint32_t pad;
} smg_plan_stats;

int smg_plan_new(smg_device* dev, const smg_plan_caps* caps, smg_plan** out);
/* Capacities holding the run of these molecules. The table-side capacities
 * are exact. The device-side ones come from `m`, the result of
 * smg_mesh_many for the same molecules, grown by the fraction `margin`; they
 * are zero when m is NULL. */
int smg_plan_caps_for(smg_resident* const* residents, const smg_mesh_params* params, size_t count,
                      const smg_mesh* m, double margin, smg_plan_caps* caps);
int smg_plan_run(smg_plan* p, smg_resident* const* residents, const smg_mesh_params* params, size_t count,
                 smg_mesh** out, smg_mesh_times* times);
/* Capacities predicted by a sizing model, for molecules no run has seen:
 * the table side exact, the device side from the model. The device marks
 * the molecules' touched bricks and runs the model on each: the expected
 * surface bricks are the sum of the predicted occupancy probabilities;
 * vertices and triangles follow from the ratios of the meshes the model
 * was trained on at the nearest spacing (vertices per surface brick per
 * b^2, triangles per vertex). Each is grown by the fraction `margin`. No mesh stage runs; a
 * plan run whose counts exceed them falls back to the exact path and grows
 * (see above), so a low estimate costs time, never a different mesh. `est`
 * (may be NULL) receives the estimates. */
typedef struct {
    uint64_t touched_bricks;
    double surface_bricks;  /* sum of sigmoid(occupancy logit) */
    double occupied_bricks; /* bricks with a positive logit */
    double vertices, triangles;
    double device_ms;
} smg_sizing_estimate;
int smg_plan_caps_predict(smg_resident* const* residents, const smg_mesh_params* params, size_t count,
                          const smg_model* model, double margin, smg_plan_caps* caps, smg_sizing_estimate* est);
int smg_plan_stats_get(const smg_plan* p, smg_plan_stats* stats);
void smg_plan_free(smg_plan* p);

#ifdef __cplusplus
}
#endif
#endif
