/*
 * octopus_engine.h - C ABI of the MI355X soft-body FEM stepping engine.
 *
 * The entry points replace the reference's GPU dynamics path, which is a C++
 * object graph driven by a script component (OctopusEngine, paths relative to
 * octopus/):
 *
 *   Cuda_FEM_Dynamic::init            include/Script/Dynamic/Cuda_FEM_Dynamic.h:18-25
 *   GPU_ParticleSystem(...)           include/GPU/GPU_ParticleSystem.h:56
 *   GPU_VBD / GPU_VBD_FEM ctors       include/GPU/VBD/GPU_VBD.h:15, include/GPU/VBD/GPU_VBD_FEM.h:62-65
 *   GPU_PBD / GPU_PBD_FEM ctors       include/GPU/PBD/GPU_PBD.h:37, include/GPU/PBD/GPU_PBD_FEM.h:19-20
 *   GPU_FEM_Explicit ctor             include/GPU/Explicit/GPU_FEM_Explicit.h:14-15
 *   GPU_Fix_Constraint / Box_Limit    include/GPU/GPU_Constraint.h:11-24
 *   step(dt, substeps)                src/Script/Dynamic/Cuda_ParticleSystem_Dynamic.cpp:17-31
 *   getters (positions, velocities..) include/GPU/GPU_ParticleSystem.h:58-65
 *
 * Conventions: plain pointers and sizes, float xyz arrays are 3*N interleaved,
 * topology is int32 S*E (element node order of the reference converters).
 * Every function returns an oe_status; on error oe_last_error() gives the
 * message of the last failing call on this thread.  Input arrays are copied;
 * the engine owns all device memory.  A scene may be used by one host thread
 * at a time.  Device work is queued on the engine's stream; the read-back
 * functions synchronise that stream.
 */
#ifndef OCTOPUS_ENGINE_H
#define OCTOPUS_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t oe_status;
enum {
    OE_OK = 0,
    OE_ERR_INVALID = 1,   /* bad argument / unsupported configuration */
    OE_ERR_HIP = 2,       /* HIP runtime error */
    OE_ERR_BUDGET = 3,    /* device memory budget exceeded */
    OE_ERR_STATE = 4,     /* call not valid in the object's current state */
    OE_ERR_INTERNAL = 5
};

/* Element ids = reference enum Element (include/Mesh/Elements.h:4-15). */
enum { OE_TETRA = 3, OE_HEXA = 6, OE_TETRA10 = 7, OE_TETRA20 = 8, OE_HEXA27 = 9 };
/* Material ids = reference enum Material (include/Dynamic/FEM/ContinuousMaterial.h:8-10). */
enum { OE_NEOHOOKE = 2, OE_STABLE_NEOHOOKE = 3 };
/* Mass_Distribution (include/Dynamic/FEM/FEM_Shape.h:693-696). */
enum { OE_MASS_UNIFORM = 0, OE_MASS_FEMSHAPE = 1 };
/* Integrator family. */
enum { OE_SOLVER_VBD = 0, OE_SOLVER_XPBD = 1, OE_SOLVER_EXPLICIT = 2 };
/* VBD_Version (include/GPU/VBD/GPU_VBD_FEM.h:14-16). */
enum { OE_VBD_BASE = 0, OE_VBD_THREADED_QUADRATURE = 1, OE_VBD_REDUCTION_SYMMETRY = 2, OE_VBD_BETTER_COLORATION = 3 };
/* XPBD_FEM_VERSION (include/GPU/PBD/GPU_PBD_FEM.h:11-13). */
enum { OE_XPBD_XPBD = 0, OE_XPBD_QUADRATURE = 1, OE_XPBD_OPTISHARED = 2, OE_XPBD_OPTIGROUP = 3 };

typedef struct oe_engine oe_engine;
typedef struct oe_scene oe_scene;
typedef struct oe_batch oe_batch;
typedef struct oe_server oe_server;

/* Scene description: the arguments of Cuda_{VBD,XPBD}_FEM_Dynamic /
 * Cuda_FEM_Explicit_Dynamic plus the mesh they are attached to. */
typedef struct oe_scene_desc {
    int32_t element;            /* OE_TETRA ... OE_HEXA27 */
    int32_t n_vertices;
    const float* xyz;           /* 3 * n_vertices */
    int32_t n_elements;
    const int32_t* topology;    /* S * n_elements */
    int32_t material;           /* OE_NEOHOOKE / OE_STABLE_NEOHOOKE */
    float young, poisson, density;
    int32_t mass_distribution;  /* OE_MASS_* */
    int32_t solver;             /* OE_SOLVER_* */
    int32_t iterations;         /* VBD iterations per substep */
    float rho;                  /* VBD Chebyshev spectral radius, 0 = off */
    float damping;              /* VBD / explicit damping */
    float global_damping;       /* XPBD velocity damping */
    int32_t vbd_version;        /* OE_VBD_* */
    int32_t target_color;       /* OE_VBD_BETTER_COLORATION target, > 0 */
    int32_t xpbd_version;       /* OE_XPBD_* */
    float gravity[3];
    uint64_t shuffle_seed;      /* 0: ascending colour order; else seeded shuffle per sweep */
} oe_scene_desc;

/* Fill a description with the reference defaults (SNH, VBD v3 semantics, g = -9.81 y). */
void oe_scene_desc_default(oe_scene_desc* d);

const char* oe_last_error(void);
const char* oe_version(void);

/* ---------------------------------------------------------------- engine */
/* budget_bytes = 0: no budget. */
oe_status oe_engine_create(int32_t device, uint64_t budget_bytes, oe_engine** out);
oe_status oe_engine_destroy(oe_engine* e);
oe_status oe_engine_synchronize(oe_engine* e);
oe_status oe_engine_memory(oe_engine* e, uint64_t* used_bytes, uint64_t* budget_bytes);

/* ------------------------------------------------------------------ scene */
oe_status oe_scene_create(oe_engine* e, const oe_scene_desc* desc, oe_scene** out);
oe_status oe_scene_destroy(oe_scene* s);
/* GPU_Fix_Constraint with an explicit vertex list. */
oe_status oe_scene_fix_vertices(oe_scene* s, const int32_t* ids, int32_t n);
/* GPU_Fix_Constraint(geometry, Plane(o, n)): vertices with dot(x - o, n) > 0. */
oe_status oe_scene_fix_plane(oe_scene* s, const float o[3], const float n[3], int32_t* n_fixed);
/* GPU_Box_Limit(pmin, pmax). */
oe_status oe_scene_box_limit(oe_scene* s, const float pmin[3], const float pmax[3]);
/* Cuda_ParticleSystem_Dynamics::update: substeps x integrator.step(dt / substeps). Asynchronous. */
oe_status oe_scene_step(oe_scene* s, float dt, int32_t substeps);
/* Capture the substep loop into a HIP graph and replay it (1) or launch eagerly (0). */
oe_status oe_scene_set_graph(oe_scene* s, int32_t enable);
/* VBD colour-pass layout: lanes per vertex (4..64, power of two) and
 * quadrature points per lane of the gather kernel, and mode = 0 (gather),
 * 1 (gather with the displacement gradients of every (element, quadrature
 * point) kept in device memory during a sweep; qpl must be 1) or 2 (element
 * tasks - one per incidence of the colour, sum-factorised gradients for
 * 27-node hexahedra - then vertex solves).  Mode 1 is a VBD option.  On an
 * explicit scene mode 2 (the default) selects the element-centric force pass
 * (one element per 32 lanes, then a per-vertex sum of the element partials)
 * and mode 0 the per-vertex gather with this tile.  No counterpart in the reference,
 * whose kernels fix the thread layout per version
 * (src/GPU/VBD/GPU_VBD_FEM.cu:712-756). */
oe_status oe_scene_set_tile(oe_scene* s, int32_t lanes, int32_t qpl, int32_t mode);
oe_status oe_scene_get_tile(oe_scene* s, int32_t* lanes, int32_t* qpl, int32_t* mode);
/* The full colour-pass layout: the gather knobs above plus, for element
 * tasks, the incidences per block (the elements one workgroup stages in LDS;
 * 0 = the kernel's maximum, 256 for tetrahedra and 8-node hexahedra, 35 for
 * 27-node hexahedra, where 14, 21, 28 and 35 select the throughput kernel with
 * 9-lane tasks and 8, 16 and 32 the one with 27-lane tasks) and the
 * throughput-form block count below which a colour runs in the kernels'
 * latency form (default 512).  No counterpart in
 * the reference (src/GPU/VBD/GPU_VBD_FEM.cu:619-650 fixes one layout per
 * kernel version). */
typedef struct oe_tile {
    int32_t lanes;
    int32_t qpl;
    int32_t mode;
    int32_t inc_per_block;
    int64_t latency_blocks;
} oe_tile;
oe_status oe_scene_set_tile_ex(oe_scene* s, const oe_tile* t);
oe_status oe_scene_get_tile_ex(oe_scene* s, oe_tile* t);

/* Tile autotuner (no counterpart in the reference).  Times every layout the
 * kernels accept for the scene's element type on the device (one colour sweep
 * each, with HIP events on the engine's stream; the state the sweeps move is
 * saved and put back), applies the fastest and caches it per (device,
 * element type, floor(log2(vertices)), scene or batch) for the rest of the
 * process.  A cached layout is applied without timing unless force != 0.
 * VBD and explicit scenes; not while the scene belongs to a batch or a server
 * (OE_ERR_STATE).  Outputs may be NULL: the chosen layout, its time per sweep
 * (ms) and the number of layouts timed (0 = taken from the cache). */
oe_status oe_scene_autotune(oe_scene* s, int32_t force, oe_tile* chosen, double* sweep_ms, int32_t* candidates);
/* When enabled, oe_scene_create tunes every VBD / explicit scene it creates on
 * this engine (cached layouts apply without timing). */
oe_status oe_engine_set_autotune(oe_engine* e, int32_t enable);
/* Entries of the autotuner's cache (all devices): capacity in *n, count out;
 * per entry device, element, size bucket, form (0 scene, 1 batch), layout,
 * time (ms per sweep, or per batched step of that element type's colour
 * passes) and candidates timed.  Any array may be NULL. */
oe_status oe_autotune_cache(int32_t* n, int32_t* device, int32_t* element, int32_t* bucket, int32_t* form,
                            oe_tile* tiles, double* ms, int32_t* candidates);
oe_status oe_autotune_clear(void);

/* Element record layout chosen at creation: affine = 1 when every element is
 * an affine image of the reference element and its J^-1, |det J| are stored
 * once per element; 0 for one record per (element, quadrature point) as in
 * the reference's GPU_FEM (src/GPU/GPU_FEM.cu:176-214). */
oe_status oe_scene_get_layout(oe_scene* s, int32_t* affine);

/* Counts: vertices, elements, nodes per element, quadrature points, colours. */
oe_status oe_scene_info(oe_scene* s, int32_t* n_vertices, int32_t* n_elements, int32_t* nodes_per_element,
                        int32_t* quadrature, int32_t* n_colors, uint64_t* device_bytes);
/* Read-back (synchronising).  Arrays of 3*N floats / N values. */
oe_status oe_scene_get_positions(oe_scene* s, float* xyz);
oe_status oe_scene_get_last_positions(oe_scene* s, float* xyz);
oe_status oe_scene_get_init_positions(oe_scene* s, float* xyz);
oe_status oe_scene_get_velocities(oe_scene* s, float* xyz);
oe_status oe_scene_get_forces(oe_scene* s, float* xyz);
oe_status oe_scene_get_masses(oe_scene* s, float* m);
oe_status oe_scene_get_masks(oe_scene* s, int32_t* mask);
/* Colour of every vertex (VBD) or element (XPBD). */
oe_status oe_scene_get_colors(oe_scene* s, int32_t* colors);
/* State export / import (p, last_p, v; any pointer may be NULL on import). */
oe_status oe_scene_set_state(oe_scene* s, const float* p, const float* last_p, const float* v);
/* Gather-kernel timing with HIP events around every colour launch. */
oe_status oe_scene_set_timing(oe_scene* s, int32_t enable);
oe_status oe_scene_get_timing(oe_scene* s, double* total_ms, int64_t* launches);

/* ------------------------------------------------- batched stepping (many scenes) */
/* Many VBD scenes stepped together: one launch per global colour index and
 * element type covers every scene of the batch (the reference steps scene after
 * scene, src/Core/Engine.cpp:15-21).  Each scene keeps its own results exactly
 * as if stepped alone.  A batch holds no copy of the scenes' positions and
 * velocities, but copies their fixed sets, box limits and task tables when it
 * is created: while a scene belongs to a batch its constraint and tile changes
 * return OE_ERR_STATE.  A scene appears once per batch and cannot be held by a
 * server; destroying a scene makes its batches refuse to step (OE_ERR_STATE). */
oe_status oe_batch_create(oe_engine* e, oe_scene* const* scenes, int32_t n, oe_batch** out);
oe_status oe_batch_destroy(oe_batch* b);
oe_status oe_batch_step(oe_batch* b, float dt, int32_t substeps);
oe_status oe_batch_set_graph(oe_batch* b, int32_t enable);
/* Streams of a batch step: 2 (default) runs the 27-node hexahedra's colour
 * passes on a second stream beside the other element types' (the two chains
 * are independent within an iteration and fill each other's idle issue
 * slots), 1 runs every pass on the engine's stream.  The results are the same
 * either way.  No counterpart in the reference. */
oe_status oe_batch_set_streams(oe_batch* b, int32_t n);
oe_status oe_batch_set_tile(oe_batch* b, int32_t element, int32_t lanes, int32_t qpl, int32_t mode);
/* Full layout per element type (latency_blocks has no effect: batches run the
 * throughput form); get returns the layout in use. */
oe_status oe_batch_set_tile_ex(oe_batch* b, int32_t element, const oe_tile* t);
oe_status oe_batch_get_tile_ex(oe_batch* b, int32_t element, oe_tile* t);
/* Autotune every element type of the batch: the candidates are timed on the
 * batch itself (the colour-pass launches of that type in one batched step,
 * HIP events); the scenes' state is saved and put back.  Cached per (device,
 * element type, floor(log2(vertices of that type)), batch) as above. */
oe_status oe_batch_autotune(oe_batch* b, int32_t force);
oe_status oe_batch_set_timing(oe_batch* b, int32_t enable);
oe_status oe_batch_get_timing(oe_batch* b, double* gather_ms, int64_t* launches);
/* Gather launches and vertex-updates per substep; algorithmic gather bytes per substep. */
oe_status oe_batch_info(oe_batch* b, int32_t* gather_launches, int64_t* vertex_updates, double* gather_bytes);
/* The same two, restricted to the colour-pass launches of one element type
 * (their own kernels): event-timed milliseconds and launch count since the
 * last read for that type, and its share of the per-substep figures. */
oe_status oe_batch_get_timing_element(oe_batch* b, int32_t element, double* gather_ms, int64_t* launches);
oe_status oe_batch_info_element(oe_batch* b, int32_t element, int32_t* gather_launches, int64_t* vertex_updates,
                                double* gather_bytes);

/* ------------------------------------------------- serving (continuous batching) */
/* A request loop over HBM-resident VBD scenes (no counterpart in the
 * reference, which steps every scene once per frame).  A request advances one
 * scene by step(dt, substeps).  A worker thread batches the oldest pending
 * requests whose scenes are idle (one request per scene per batch, at most
 * max_batch scenes) into one batched step on one of `streams` streams, so
 * requests that arrive while others are on the GPU are admitted as soon as a
 * stream and their scene are free and batches overlap on the device.  Scenes
 * are admitted against the engine's memory budget when they are created
 * (oe_scene_create returns OE_ERR_BUDGET beyond it).  Each stream has its own
 * worker thread; a tick holds scenes of one element type.  Streams only run
 * concurrently on hardware queues of their own: keep streams + 2 (the
 * engine's stream and the null stream) <= GPU_MAX_HW_QUEUES, whose HIP default
 * is 4, set in the environment before the process first uses HIP. */
oe_status oe_server_create(oe_engine* e, int32_t max_batch, int32_t streams, oe_server** out);
oe_status oe_server_destroy(oe_server* sv);
/* Register a scene; returns its slot id.  A scene joins at most one server,
 * once, and not while it belongs to a batch (OE_ERR_INVALID / OE_ERR_STATE).
 * While it is held: oe_scene_step, oe_scene_set_state and constraint or tile
 * changes return OE_ERR_STATE; reads (positions, velocities, diagnostics,
 * residual, VTK) first wait for every request submitted for the scene;
 * oe_scene_destroy waits for them and releases the slot. */
oe_status oe_server_add_scene(oe_server* sv, oe_scene* s, int32_t* slot);
/* Layout of one element type's colour passes in every tick (default: the
 * scene defaults, with 14 incidences per block for 27-node hexahedra).
 * Refused (OE_ERR_STATE) while requests are pending or in flight; get returns
 * the layout in use.  No counterpart in the reference. */
oe_status oe_server_set_tile_ex(oe_server* sv, int32_t element, const oe_tile* t);
oe_status oe_server_get_tile_ex(oe_server* sv, int32_t element, oe_tile* t);
/* Thread-safe; records the arrival time of the request.  Requests of one
 * scene run in submission order. */
oe_status oe_server_submit(oe_server* sv, int32_t slot, float dt, int32_t substeps, uint64_t* ticket);
/* Wait until every submitted request is complete. */
oe_status oe_server_drain(oe_server* sv);
/* Wait until the request with this ticket is complete (thread-safe). */
oe_status oe_server_wait(oe_server* sv, uint64_t ticket);
/* *done = 1 once the request with this ticket is complete, else 0 (thread-safe). */
oe_status oe_server_poll(oe_server* sv, uint64_t ticket, int32_t* done);
/* Since creation: completed requests, latency percentiles (ms, arrival to
 * completion), vertex-updates, batches and the mean batch size. */
oe_status oe_server_stats(oe_server* sv, int64_t* completed, double* p50_ms, double* p99_ms, double* max_ms,
                          int64_t* vertex_updates, int64_t* batches, double* mean_batch);
/* Per-request latencies in completion order (ms), up to cap entries. */
oe_status oe_server_latencies(oe_server* sv, double* out, int64_t cap, int64_t* n);
/* The slot of each request in the same order (for latency by scene). */
oe_status oe_server_latency_slots(oe_server* sv, int32_t* out, int64_t cap, int64_t* n);

/* ------------------------------------------------- host-side (no GPU needed) */
/* Beam of nx*ny*nz cells of size sx*sy*sz; pass NULL arrays to query sizes. */
oe_status oe_host_beam_mesh(int32_t element, int32_t nx, int32_t ny, int32_t nz, float sx, float sy, float sz,
                            int32_t* n_vertices, int32_t* n_elements, float* xyz, int32_t* topology);
/* Nodes per element and quadrature points; dN at the quadrature points (Q*S*3). */
oe_status oe_host_shape(int32_t element, int32_t* S, int32_t* Q, double* dN, double* weights, double* qpoints);
/* DSAT colouring of the primal graph (VBD); target_color > 0 applies ColorReduction. */
oe_status oe_host_color_primal(int32_t element, const int32_t* topology, int32_t n_elements, int32_t n_vertices,
                               int32_t target_color, int32_t* colors, int32_t* n_colors, int64_t* conflicts);
/* Greedy smallest-last colouring of the element (vertex-dual) graph (XPBD). */
oe_status oe_host_color_dual(int32_t element, const int32_t* topology, int32_t n_elements, int32_t n_vertices,
                             int32_t* colors, int32_t* n_colors, int64_t* conflicts);
/* JX_inv (column-major) + V per (element, quadrature point): 10 floats each. */
oe_status oe_host_fem_constants(int32_t element, const float* xyz, int32_t n_vertices, const int32_t* topology,
                                int32_t n_elements, float* jxv);
/* Per-element records of an affine mesh (every element X_k = a + B xi_k up to
 * coordinate rounding): records = 10 floats per element (B^-1 column-major,
 * |det B|), weights = the Q quadrature weights, V_q = |det B| weights[q].
 * *affine = 0 (outputs untouched) when some element is curved.  No counterpart
 * in the reference, which stores one record per (element, quadrature point). */
oe_status oe_host_affine_records(int32_t element, const float* xyz, int32_t n_vertices, const int32_t* topology,
                                 int32_t n_elements, float* records, float* weights, int32_t* affine);
oe_status oe_host_masses(int32_t element, const float* xyz, int32_t n_vertices, const int32_t* topology,
                         int32_t n_elements, float density, int32_t distribution, float* masses);

/* ------------------------------------------------------------ diagnostics */
/* Per-element values of the current state, n_elements floats each (any output
 * may be NULL; synchronising).  Replaces GPU_FEM::get_stress / get_volume /
 * get_volume_diff / get_inverted (src/GPU/GPU_FEM.cu:117-152):
 *   stress = sum_q vonMises(cauchy(F_q)) V_q, volume = sum_q |det F_q| V_q,
 *   volume_diff = volume - sum_q V_q, inverted = 1 if sum_q det F_q w_q <= -1e-6. */
oe_status oe_scene_element_diagnostics(oe_scene* s, float* stress, float* volume, float* volume_diff,
                                       float* inverted);
/* Residual force of each free vertex's local VBD problem at the current state
 * (3 * n_vertices floats; fixed vertices 0).  Replaces GPU_VBD_FEM::get_forces
 * (src/GPU/VBD/GPU_VBD_FEM.cu:482-559). */
oe_status oe_scene_residual(oe_scene* s, float dt, float* r);
/* The same residual with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its kernel writes into a static device array of 3 floats per vertex
 * (scenes of at most 524288 vertices, else OE_ERR_INVALID; 6 MiB held on each
 * device the library runs on until the process ends) that every call on one
 * device shares: calls that overlap on one device race.  Results equal
 * oe_scene_residual's.  Synchronising. */
oe_status oe_scene_residual_scratch(oe_scene* s, float dt, float* r);

/* ------------------------------------------------------------ mesh files */
/* Legacy-VTK ASCII unstructured grids (.vtk) and Gmsh 2.2 ASCII (.msh), as the
 * reference's VTK_Loader / Msh_Loader read them
 * (src/Mesh/Converter/MeshLoader.cpp:27-237).  counts[element id] (10
 * entries) = number of elements of that type; topologies are in the engine's
 * (the reference's) node order. */
typedef struct oe_mesh oe_mesh;
oe_status oe_mesh_load(const char* path, oe_mesh** out);
oe_status oe_mesh_info(oe_mesh* m, int32_t* n_vertices, int32_t* counts);
oe_status oe_mesh_vertices(oe_mesh* m, float* xyz);
oe_status oe_mesh_topology(oe_mesh* m, int32_t element, int32_t* topology);
oe_status oe_mesh_destroy(oe_mesh* m);

/* ------------------------------------------------------------ VTK output */
/* Legacy-VTK 5.1 ASCII unstructured grid in the layout of the reference's
 * VTK_Formater (src/Mesh/Converter/VTK_Formater.cpp): open, mesh (POINTS,
 * CELLS/OFFSETS/CONNECTIVITY, CELL_TYPES; Hexa27 cells in VTK node order),
 * then point_data / cell_data blocks of scalars (n floats) and vectors (3n
 * floats), close (flushes and frees the handle). */
typedef struct oe_vtk oe_vtk;
oe_status oe_vtk_open(const char* path, oe_vtk** out);
oe_status oe_vtk_mesh(oe_vtk* v, int32_t element, const float* xyz, int32_t n_vertices, const int32_t* topology,
                      int32_t n_elements);
oe_status oe_vtk_point_data(oe_vtk* v);
oe_status oe_vtk_cell_data(oe_vtk* v);
oe_status oe_vtk_scalars(oe_vtk* v, const char* name, const float* data, int32_t n);
oe_status oe_vtk_vectors(oe_vtk* v, const char* name, const float* data, int32_t n);
oe_status oe_vtk_close(oe_vtk* v);
/* Scene snapshot in the layout of FEM_VTK_Recorder::save
 * (src/Script/Record/Recorder.cpp:117-171): the rest geometry; point data
 * "weights" (vertex masses), "u_norm" (|u|), "u" (displacement),
 * "smooth_stress" (sum of the incident elements' stress / vertex mass); cell
 * data "stress", "volume", "volume_diff", "inverted" (the element diagnostics
 * above).  The reference's recorder copies its "inverted" array onto itself
 * (:137-141) and so writes zeros; this writes the flags.  Synchronising. */
oe_status oe_scene_save_vtk(oe_scene* s, const char* path);

/* ------------------------------------------------------------ JSON record */
/* The reference's DataRecorder file (src/Script/Record/DataRecorder.cpp:16-38):
 * {"name" : value,...} in the order the entries are added, each value written
 * as its recorder writes it (std::ostream default formatting):
 *   time      TimeRecorder      (include/Script/Record/Recorder.h:37-47)
 *   mesh      MeshRecorder      (src/Script/Record/Recorder.cpp:13-24); counts[element id], 10 entries
 *   material  FEM_Dynamic_Recorder (xpbd = 0: sub_step) or XPBD_FEM_Dynamic_Recorder
 *             (xpbd = 1: iteration, sub_itetrations) (:27-49), both named "XPBD_FEM_Dynamic"
 *   string    the VTK recorders' file entries (:95-97, 113-115, 173-175)
 *   flexion   FEM_Flexion_error_recorder (:210-219)
 *   rotation  FEM_Torsion_error_recorder (:317-341)
 * close writes the file and frees the handle; text copies the JSON text into
 * buf (capacity cap bytes, NUL-terminated) and sets *len to its length. */
typedef struct oe_record oe_record;
oe_status oe_record_open(const char* path, oe_record** out);
oe_status oe_record_time(oe_record* r, int32_t frame, float time, float dt, float fixed_time, float fixed_dt);
oe_status oe_record_mesh(oe_record* r, const int32_t* counts, int32_t n_vertices);
oe_status oe_record_material(oe_record* r, int32_t xpbd, int32_t material, float poisson, float young,
                             int32_t iterations, int32_t sub_iterations, float density);
oe_status oe_record_string(oe_record* r, const char* name, const char* value);
oe_status oe_record_flexion(oe_record* r, float error, const float* p, const float* target);
oe_status oe_record_rotation(oe_record* r, float rot_max, float beam_length, const float* dist, const float* angles,
                             int32_t n);
oe_status oe_record_text(oe_record* r, char* buf, int64_t cap, int64_t* len);
oe_status oe_record_close(oe_record* r);

/* Error measures of the reference's recorders, on host arrays (3 floats per
 * vertex).  Flexion (FEM_Flexion_error_recorder::init, Recorder.cpp:190-208):
 * the first vertex whose rest position is within squared distance 1e-6 of
 * follow; *error = |target - p|^2, p its current position, *vertex its index
 * (-1 and no outputs when there is none).  Rotation
 * (FEM_Torsion_error_recorder::init + compute_errors, Recorder.cpp:221-315):
 * samples_per_edge samples on every element edge lying on y = z = 0 in the rest
 * state; dist = current x / beam_length, angle (degrees) between the rest and
 * current (y, z) directions from the axis (0.5, 0.5); sorted by dist, runs
 * closer than 1e-4 merged.  *n: capacity in, count out (OE_ERR_INVALID when
 * the capacity is short). */
oe_status oe_flexion_error(const float* rest, const float* pos, int32_t n_vertices, const float* follow,
                           const float* target, float* error, float* p, int32_t* vertex);
oe_status oe_rotation_error(int32_t element, const int32_t* topology, int32_t n_elements, const float* rest,
                            const float* pos, int32_t n_vertices, int32_t samples_per_edge, float beam_length,
                            float* dist, float* angles, int32_t* n);

#ifdef __cplusplus
}
#endif

#endif /* OCTOPUS_ENGINE_H */
