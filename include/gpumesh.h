/* gpumesh.h -- C ABI of libgpumesh_hip.so, the MI355X vertex-relocation
 * smoothing library.
 *
 * Plain C types only (pointers, sizes, status codes).  A mesh handle owns a
 * device-resident copy of the mesh (positions in FP64 and FP32, connectivity,
 * patch lists, colour groups) and one HIP stream; calls on one handle are
 * serialised by the caller (a handle is not reentrant), calls on different
 * handles are independent.
 *
 * Numbering: gm_mesh_create renumbers vertices exactly as the reference's
 * NodeGroups::clusterNodes does (fixed vertices first, then colour group by
 * colour group).  Every array exchanged after creation (positions, per-vertex
 * counts, connectivity) uses the NEW numbering; gm_mesh_permutation returns
 * old id -> new id.
 *
 * Every function returns GM_OK (0) or a GM_ERR_* code; gm_last_error() gives
 * a message for the calling thread.
 */
#ifndef GPUMESH_H
#define GPUMESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gm_mesh gm_mesh;

enum gm_status
{
    GM_OK = 0,
    GM_ERR_ARG = 1,          /* invalid argument */
    GM_ERR_HIP = 2,          /* HIP runtime error (no device, launch failure...) */
    GM_ERR_UNSUPPORTED = 3,  /* valid request this build does not implement */
    GM_ERR_NOMEM = 4         /* device allocation failed */
};

enum gm_smoother { GM_NELDER_MEAD = 0, GM_GRADIENT_DESCENT = 1, GM_SPAWN_SEARCH = 2, GM_GETME = 3 };
enum gm_evaluator { GM_MEAN_RATIO = 0, GM_METRIC_CONFORMITY = 1 };
/* GM_ANALYTIC_SHELL: the spherical-shell field of the anisotropic hexahedral
 * workload; GM_ANALYTIC: the reference's AnalyticSampler field (a rotated
 * 1-D sinusoidal anisotropy, Samplers/AbstractSampler.cpp:190-208; aspect
 * ratio from gm_mesh_set_aspect_ratio). */
enum gm_sampler { GM_UNIFORM = 0, GM_ANALYTIC_SHELL = 1, GM_TEXTURE = 2, GM_ANALYTIC = 3 };
/* GM_FP64: every pass in FP64 with the reference's operation order (bit-exact
 * with an FP64 CPU build).  GM_FP32: every pass in FP32.  GM_FP32_POLISH: FP32
 * passes, then the last gm_set_polish_passes() passes (default 1) in FP64 on
 * the mixed-precision kernels.  GM_FP64_FAST: every pass on those FP64
 * kernels (FP64 arithmetic in another operation order, d^(-2/3) from a
 * hardware seed refined to FP64).  Mesh-quality reductions are FP64 in every
 * mode. */
enum gm_precision { GM_FP64 = 0, GM_FP32 = 1, GM_FP32_POLISH = 2, GM_FP64_FAST = 3 };
/* Kernel shape: one vertex per lane (plain), or one vertex per group of
 * lanes holding its patch elements (wave). */
enum gm_kernel { GM_KERNEL_PLAIN = 0, GM_KERNEL_WAVE = 1 };

const char* gm_last_error(void);
const char* gm_version(void);

/* Number of visible HIP devices (0 without a GPU). */
int gm_device_count(int* count);

/* Builds topology and colour groups on the host (bit-exact with the
 * reference's Mesh::compileTopology, DataStructures/Mesh.cpp:213-284, and
 * NodeGroups::build, DataStructures/NodeGroups.cpp:104-122) and uploads the
 * mesh to the current device.  pos: 3*nverts doubles; fixed: nverts flags
 * (non-zero = vertex fixed in place, the reference's MeshTopo(fixedPosition),
 * Mesh.cpp:144-147); tets: 4*ntets ids; hexs: 8*nhexs ids (v0..v3 bottom,
 * v4..v7 top).  Replaces Mesh::updateCudaTopology/updateCudaVertices
 * (GpuMesh.cpp:227-276 -> resources/cuda/Mesh.cu:74-279). */
int gm_mesh_create(int64_t nverts, const double* pos, const int32_t* fixed,
                   int64_t ntets, const uint32_t* tets, int64_t nhexs, const uint32_t* hexs,
                   gm_mesh** out);
/* Same as gm_mesh_create but builds the topology and colour groups on the
 * host only (no device copy, no GPU needed): the topology queries below work,
 * calls that need the device return GM_ERR_ARG. */
int gm_mesh_create_host(int64_t nverts, const double* pos, const int32_t* fixed,
                        int64_t ntets, const uint32_t* tets, int64_t nhexs, const uint32_t* hexs,
                        gm_mesh** out);
/* A mesh in the caller's own numbering with the caller's colour groups,
 * the form in which the reference already holds them: group g is
 * group_nodes[group_off[g] .. group_off[g+1]) (group_off[0] = 0, ngroups + 1
 * entries), the concatenation of NodeGroups::parallelGroups()[g]
 * .undispatchedNodes (DataStructures/NodeGroups.cpp:452-619; the GPU buffer
 * of NodeGroups::gpuGroupsBuffer, :621-693, when no vertex slides on the
 * boundary).  Fixed vertices belong to no group; every other vertex to
 * exactly one; no two vertices of a group may share an element -- all
 * checked on the host (GM_ERR_ARG naming the offending vertices).  Sweeps
 * visit the groups in order.  No renumbering: positions, elements and counts
 * stay in the caller's numbering and gm_mesh_permutation returns the
 * identity.  device_copy = 0 builds the host topology only (as
 * gm_mesh_create_host).  Replaces Mesh::updateCudaTopology and
 * updateCudaGroupMembers (DataStructures/GpuMesh.cpp:14-25, 227-276). */
int gm_mesh_create_grouped(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets,
                           const uint32_t* tets, int64_t nhexs, const uint32_t* hexs, int64_t ngroups,
                           const int64_t* group_off, const uint32_t* group_nodes, int device_copy, gm_mesh** out);
/* ---- Meshes with sliding boundary vertices.
 *
 * The reference's boundaries (Boundaries/, e.g. BoxBoundary on [-1,1]^3,
 * Boundaries/BoxBoundary.cpp:7-36) let vertices slide on faces and edges:
 * every trial position of such a vertex is projected onto its constraint
 * (MeshTopo::snapToBoundary, DataStructures/Mesh.h:256-267).  A constraint
 * is a plane through a with normal b (PlaneConstraint::operator(),
 * Boundaries/Constraints/FaceConstraint.cpp:180-185) or a segment from a to b
 * (SegmentConstraint::operator(), EdgeConstraint.cpp:185-199); a fixed
 * vertex is a point constraint (the `fixed` flag).  Vertex types and the
 * numbering follow NodeGroups (NodeGroups.cpp:124-450): a sliding vertex is
 * a boundary vertex, a free vertex of one of its elements a subsurface
 * vertex; vertex-wise sweeps run colour by colour, GETMe's vertex steps in
 * NodeGroups' serial order (boundary, subsurface, interior). */
enum gm_constraint_kind { GM_CONSTRAINT_PLANE = 1, GM_CONSTRAINT_SEGMENT = 2 };
typedef struct gm_constraint
{
    int kind;        /* gm_constraint_kind */
    double a[3];     /* plane: a point on it; segment: first end */
    double b[3];     /* plane: its normal (any length); segment: second end */
} gm_constraint;

typedef struct gm_mesh_desc
{
    int64_t nverts;
    const double* pos;            /* 3 nverts */
    const int32_t* fixed;         /* nverts flags */
    int64_t ntets;
    const uint32_t* tets;         /* 4 ntets */
    int64_t nhexs;
    const uint32_t* hexs;         /* 8 nhexs */
    /* Sliding constraints (optional): constraint[v] = index into
     * constraints[0 .. nconstraints), or -1; a fixed vertex takes -1. */
    const int32_t* constraint;
    int64_t nconstraints;
    const gm_constraint* constraints;
    /* Colour groups (optional): ngroups < 0 colours as gm_mesh_create
     * does (renumbering); ngroups >= 0 takes the caller's numbering and
     * groups as gm_mesh_create_grouped does. */
    int64_t ngroups;
    const int64_t* group_off;
    const uint32_t* group_nodes;
    int device_copy;              /* 0: host topology only */
    /* Prisms (MeshPri, DataStructures/Mesh.h): 6 vertices each, the first
     * triangle then the one above it; npris = 0 for none. */
    int64_t npris;
    const uint32_t* pris;         /* 6 npris */
} gm_mesh_desc;

/* gm_mesh_create / gm_mesh_create_grouped from a description, with sliding
 * constraints.  A zero-length normal or segment is refused (GM_ERR_ARG). */
int gm_mesh_create_desc(const gm_mesh_desc* desc, gm_mesh** out);

int gm_mesh_destroy(gm_mesh* mesh);

/* info[8] = {nverts, ntets, nhexs, nmovable, ngroups, sum of patch sizes,
 *            sum of neighbour counts, largest patch}. */
int gm_mesh_info(const gm_mesh* mesh, int64_t* info);
/* perm[nverts]: old id -> new id (NodeGroups::determinePositions, NodeGroups.cpp:447-449). */
int gm_mesh_permutation(const gm_mesh* mesh, int64_t* perm);
/* group_off[ngroups+1], group_nodes[nmovable]: colour groups (NodeGroups::ParallelGroup). */
int gm_mesh_groups(const gm_mesh* mesh, int64_t* group_off, uint32_t* group_nodes);
/* Renumbered connectivity: tets[4*ntets], hexs[8*nhexs] (either may be NULL). */
int gm_mesh_elements(const gm_mesh* mesh, uint32_t* tets, uint32_t* hexs);
/* Renumbered prisms: *npris (may be NULL) and pris[6*npris] (may be NULL).
 * Replaces reading Mesh::pris after NodeGroups::clusterNodes
 * (DataStructures/NodeGroups.cpp:452-619). */
int gm_mesh_prisms(const gm_mesh* mesh, int64_t* npris, uint32_t* pris);
/* Patch lists and neighbour lists (MeshTopo::neighborElems / neighborVerts):
 * eoff[nverts+1], eref[sum patch] = (elem id << 4) | (is_record << 3) | local index,
 * where a record id below nhexs is a hexahedron and nhexs + p is prism p,
 * noff[nverts+1], nbr[sum neighbours]. */
int gm_mesh_topology(const gm_mesh* mesh, uint32_t* eoff, uint32_t* eref, uint32_t* noff, uint32_t* nbr);

/* FP64 positions, new numbering (replaces fetchCudaVerts, resources/cuda/Mesh.cu:56-72);
 * on a mesh from gm_mesh_create_host, its input positions. */
int gm_mesh_get_positions(gm_mesh* mesh, double* pos);
/* Replaces updateCudaVerts (resources/cuda/Mesh.cu:74-90). */
int gm_mesh_set_positions(gm_mesh* mesh, const double* pos);

/* Crew: evaluator, sampler and its scaling (the measurer follows the sampler
 * as in MeshCrew::setSampler, DataStructures/MeshCrew.cpp:93-103: uniform ->
 * metric-free, analytic/texture -> metric-wise).  Texture arguments are used
 * only for GM_TEXTURE: tsize[3] cells, tmin[3]/text[3] grid box,
 * texels[6*cells] FP32 (m00 m01 m02 m11 m12 m22, cell (i,j,k) at
 * (k*ny + j)*nx + i).  Replaces installCuda*Evaluator/Measurer/Sampler,
 * setCudaMetricScaling* (Samplers/AbstractSampler.cpp:11-15, 103-111) and
 * updateCudaSamplerTextures (Samplers/TextureSampler.cpp:75-80).  A mesh
 * from gm_mesh_create_host takes every sampler but GM_TEXTURE (its crew
 * serves gm_mesh_restructure). */
int gm_mesh_set_crew(gm_mesh* mesh, int evaluator, int sampler, double scaling,
                     const int32_t* tsize, const double* tmin, const double* text, const float* texels);

/* Aspect ratio of the GM_ANALYTIC field (AbstractSampler::setAspectRatio,
 * Samplers/AbstractSampler.cpp:56-59; default 1), kept across
 * gm_mesh_set_crew calls.  Replaces setCudaMetricAspectRatio
 * (AbstractSampler.cpp:14, 103-111). */
int gm_mesh_set_aspect_ratio(gm_mesh* mesh, double aspect_ratio);

/* Mesh quality in FP64 (AbstractEvaluator::evaluateMeshQuality*,
 * Evaluators/AbstractEvaluator.cpp:384-453, 574-610; replaces
 * evaluateCudaMeshQuality).  out3 = {min, sum of 1/q (+inf if any q <= 0),
 * element count}; hist20 = 20-bucket histogram; elem_qual (may be NULL) =
 * per-element qualities, tets then hexes.  Positions are read from the FP64
 * copy. */
int gm_mesh_evaluate(gm_mesh* mesh, double* out3, int32_t* hist20, double* elem_qual);

/* Runs `passes` smoothing sweeps (AbstractSmoother::smoothMesh with
 * topology operations and autopilot off, Smoothers/AbstractSmoother.cpp:57-211;
 * replaces smoothCudaVertices / smoothCudaSpawnVertices / smoothCudaElements +
 * updateCudaSmoothedElementsVertices).  Each sweep visits the colour groups in
 * order.  stats[3*(passes+1)] = per record {min, harmonic mean, sum 1/q},
 * record 0 before the first sweep; hist[20*(passes+1)]; eval_counts[nverts]
 * (may be NULL) = patch evaluations per vertex, accumulated over the run.
 * spawn_coeff: spawn-search lattice scale relative to the local element size
 * (the reference uses 1/20, SpawnSearchSmoother.cpp:14). */
int gm_mesh_smooth(gm_mesh* mesh, int smoother, int passes, int precision, int kernel,
                   double spawn_coeff, double* stats, int32_t* hist, int32_t* eval_counts);

/* Pass schedule of AbstractSmoother::evaluateMeshQuality
 * (Smoothers/AbstractSmoother.cpp:103-211; defaults Schedule.cpp:4-13:
 * autopilot off, thresholds 0.001 and 0.0).  autopilot == 0: exactly
 * max_passes sweeps (the schedule's relocationPassCount).  autopilot != 0:
 * after pass record 0 a sweep runs if max_passes > 0; after record k >= 1 the
 * next sweep runs while
 *   mean_k - mean_{k-1} > min_qual_threshold  ||  min_k - min_{k-1} > qual_mean_threshold
 * -- the reference compares the harmonic-mean gain with minQualThreshold and
 * the minimum gain with qualMeanThreshold (AbstractSmoother.cpp:186-193), and
 * these fields keep its names.  The reference's autopilot has no pass limit;
 * here max_passes bounds it.  The decision is made on the device after each
 * pass's quality reduction: the host queues max_passes passes without
 * waiting, and the passes after the stop do no work. */
typedef struct gm_schedule
{
    int autopilot;
    double min_qual_threshold;
    double qual_mean_threshold;
} gm_schedule;

/* gm_mesh_smooth under a schedule.  stats/hist are sized for max_passes + 1
 * records; the first *passes_run + 1 are written.  The autopilot takes
 * precision GM_FP64 or GM_FP32 (GM_FP32_POLISH needs a known last pass:
 * GM_ERR_UNSUPPORTED).  Replaces the pass loop of
 * AbstractVertexWiseSmoother::smoothMesh{Serial,Thread}
 * (AbstractVertexWiseSmoother.cpp:53-152) for a Schedule with topology
 * operations off. */
int gm_mesh_smooth_schedule(gm_mesh* mesh, int smoother, int max_passes, const gm_schedule* schedule,
                            int precision, int kernel, double spawn_coeff, double* stats, int32_t* hist,
                            int32_t* eval_counts, int32_t* passes_run);

/* BATR topology operations on a tetrahedral mesh, the step before the
 * relocation passes of each global pass of the reference's loop: replaces
 * BatrTopologist::restructureMesh (Topologists/BatrTopologist.cpp:76-128),
 * its Mesh::compileTopology included (AbstractVertexWiseSmoother.cpp:53-80
 * calls it).  Face swaps, edge swaps and edge splits/merges run on the host
 * over the mesh's current positions, in the reference's order, with
 * topo_passes = Schedule::topoOperationPassCount, refinement_sweeps =
 * Schedule::refinementSweepCount and edge lengths measured in the crew's
 * metric against min_edge / max_edge (AbstractTopologist's bounds, 0.5 and
 * 1.5 by default).  The mesh is then rebuilt in place: new vertex and tet
 * counts (gm_mesh_info), numbering and colour groups, device copy and the
 * crew kept.  counts (may be null): face swaps, edge swaps, splits, merges,
 * passes.  A mesh with prisms or hexahedra, or without tets, is left as it
 * is (needTopologicalModifications, :67-74); vertices that slide on
 * constraints give GM_ERR_UNSUPPORTED.  A member of a live batch gives
 * GM_ERR_ARG (destroy the batch first, build a new one afterwards).  On a mesh from gm_mesh_create_host the operations run
 * on its input positions and the result stays host-only.  A non-finite
 * position gives GM_ERR_ARG. */
int gm_mesh_restructure(gm_mesh* mesh, int topo_passes, int refinement_sweeps, double min_edge, double max_edge,
                        int64_t* counts);

/* One partial sweep of a vertex-wise smoother (NM, GD, spawn search) over the
 * colour groups [group_begin, group_end), with no quality evaluation: the
 * body of one pass of AbstractVertexWiseSmoother::smoothMeshThread
 * (Smoothers/VertexWise/AbstractVertexWiseSmoother.cpp:114-138) restricted to
 * those groups.  precision: GM_FP64, GM_FP64_FAST or GM_FP32 (FP32 rounds
 * the positions once before and widens them after).  GETMe runs only over
 * every group, [0, ngroups) (its sweep starts with an element step over the
 * whole mesh); other ranges return GM_ERR_UNSUPPORTED. */
int gm_mesh_smooth_groups(gm_mesh* mesh, int smoother, int64_t group_begin, int64_t group_end, int precision,
                          int kernel, double spawn_coeff, int32_t* eval_counts);

/* Wall time of the last gm_mesh_smooth on the device (HIP events around the
 * sweeps and evaluations), in milliseconds; out[0] = total, out[1] = sweeps only. */
int gm_mesh_last_timing(const gm_mesh* mesh, double* out2);

/* Kernel classes of gm_mesh_kernel_times. */
enum gm_kernel_class
{
    GM_KT_SMOOTH_F32 = 0,      /* vertex-wise optimiser launches (one per colour group), FP32 */
    GM_KT_SMOOTH_F64 = 1,      /* same, FP64 */
    GM_KT_GETME_ELEM_F32 = 2,  /* GETMe element transforms, FP32 */
    GM_KT_GETME_ELEM_F64 = 3,
    GM_KT_GETME_VERT_F32 = 4,  /* GETMe vertex updates (one per colour group), FP32 */
    GM_KT_GETME_VERT_F64 = 5,
    GM_KT_EVALUATE = 6,        /* FP64 mesh-quality reduction (one per pass record) */
    GM_KT_COUNT = 7
};

/* Per-class device time (HIP events around each launch on the mesh's
 * stream) and launch counts of the last gm_mesh_smooth: ms[GM_KT_COUNT],
 * launches[GM_KT_COUNT] (either may be NULL). */
int gm_mesh_kernel_times(const gm_mesh* mesh, double* ms, int64_t* launches);

/* Host topology cache shared by every mesh created afterwards: meshes with
 * identical connectivity and fixed flags (differing only in positions) reuse
 * one built topology (patch lists, colour groups, renumbering), found by an
 * exact comparison of the inputs.  max_bytes bounds the host memory it holds
 * (least recently used entries go first); 0, the default, disables it. */
int gm_set_topology_cache(int64_t max_bytes);
/* out4 = {hits, misses, entries, bytes held}.  The bytes are the host
 * topologies'; while a topology is cached, the device copy of its arrays
 * (elements, patch and neighbour lists, fixed flags, constraint ids, slot
 * map: about the same size), made by the first upload of one of its meshes
 * on each device, stays resident with it, and later meshes of that
 * topology upload only their positions. */
int gm_topology_cache_stats(int64_t* out4);

/* Builds a mesh's topology into the topology cache without creating a mesh
 * (a server pre-building the shapes it will see, so that no cold build sits
 * on a job's latency path); with device_copy, also its device arrays on the
 * current device.  The positions only serve the build.  GM_ERR_ARG when the
 * cache is off, GM_ERR_NOMEM when the topology exceeds its capacity.
 * Thread-safe.  No counterpart in the reference. */
int gm_topology_prepare(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets,
                        const uint32_t* tets, int64_t nhexs, const uint32_t* hexs, int device_copy);

/* Number of FP64 passes that end a GM_FP32_POLISH run (1 to 64, default 1;
 * a run of fewer passes runs them all in FP64).  Applies to runs, batches
 * and server jobs started afterwards, process-wide. */
int gm_set_polish_passes(int passes);

/* Device memory is allocated through a caching allocator: blocks freed by
 * destroyed meshes, batches and finished serving jobs are kept (per device,
 * in size classes of at most 25 % rounding) and reused, so that a serving
 * loop neither waits on hipMalloc nor synchronises the device in hipFree.
 * max_cached_bytes bounds the free blocks kept (default 16 GiB); 0 returns
 * them all to HIP now and disables caching. */
int gm_set_device_pool(int64_t max_cached_bytes);

/* Device memory owned by a mesh handle (positions, connectivity, patch
 * lists, reduction buffers, GETMe scratch once allocated), in bytes. */
int gm_mesh_device_bytes(const gm_mesh* mesh, int64_t* bytes);

/* ---- Batches: many resident meshes smoothed by fused launches.
 *
 * The reference smooths one mesh per AbstractSmoother::smoothMesh call
 * (Smoothers/AbstractSmoother.cpp:57-211); a batch runs that pass loop for
 * n meshes at once: every colour step of every member goes into one launch
 * per element kind, the GETMe element step and the quality reductions are
 * one launch each over all members.  Members keep their own positions,
 * statistics and pass records; each member's results are bit-identical to
 * gm_mesh_smooth on that mesh alone (same kernels, same per-mesh reduction
 * tree).  Members need the same evaluator, sampler, scaling and aspect ratio
 * (under the texture sampler also the same grid and texels, compared by
 * content) and must live on one device.  A mesh may belong to
 * several batches, but runs that touch one mesh must not overlap. */
typedef struct gm_batch gm_batch;

int gm_batch_create(gm_mesh* const* meshes, int64_t n, gm_batch** out);
int gm_batch_destroy(gm_batch* batch);
/* Launch shape of `smoother`: patches_per_group in {1,2,4,8} wavefronts per
 * workgroup (spawn search: one vertex per wavefront; gradient descent: 8;
 * GETMe vertex update: 64), tile_elems in {64,128,256} LDS elements per
 * wavefront for the candidate-parallel spawn-search and gradient-descent
 * kernels (gradient descent gives each of its 8 vertices tile_elems/8); a
 * work list whose largest patch does not fit the tile runs on the wave kernel
 * instead.  Nelder-Mead ignores it.  Until set (or autotuned) a batch uses
 * one wavefront per workgroup (GETMe: 4) and, for each work list, the
 * smallest tile that holds the list's largest patch -- the choice a member
 * makes alone. */
int gm_batch_set_shape(gm_batch* batch, int smoother, int patches_per_group, int tile_elems);
/* gm_mesh_smooth over every member.  use_graph != 0 captures one pass body
 * (all sweeps + the quality reduction) in a HIP graph on first use and
 * replays it.  stats[n][3*(passes+1)], hist[n][20*(passes+1)] (either may be
 * NULL), member order. */
int gm_batch_smooth(gm_batch* batch, int smoother, int passes, int precision, int kernel, double spawn_coeff,
                    int use_graph, double* stats, int32_t* hist);
/* gm_mesh_smooth_schedule over every member: each member stops on its own
 * records; passes_run[n] (may be NULL), stats/hist strided by max_passes + 1
 * records per member. */
int gm_batch_smooth_schedule(gm_batch* batch, int smoother, int max_passes, const gm_schedule* schedule,
                             int precision, int kernel, double spawn_coeff, int use_graph, double* stats,
                             int32_t* hist, int32_t* passes_run);
/* out2 = {total device ms of the last gm_batch_smooth, sweep ms (0 with graphs)}. */
int gm_batch_last_timing(const gm_batch* batch, double* out2);
/* Per-class device time of the last gm_batch_smooth (not recorded with graphs). */
int gm_batch_kernel_times(const gm_batch* batch, double* ms, int64_t* launches);
/* Times `passes` sweeps of `smoother` (spawn search, gradient descent or
 * GETMe) for every launch shape of the 4 x 3 grid patches_per_group x
 * tile_elems (times[12], summed sweep kernel ms; -1 where a shape does not
 * fit every list, and for GETMe where tile_elems != 64, which it does not
 * use), restores the positions after each trial, keeps the fastest shape for
 * `smoother` and returns it in best[2] (tile 0 if no shape fit: the default
 * per-list choice stays). */
int gm_batch_autotune(gm_batch* batch, int smoother, int precision, int passes, int32_t* best, double* times);

/* ---- Serving: a request loop over resident meshes.
 *
 * Jobs (a mesh, a smoother, a pass count, a precision mode) are submitted
 * from any thread and return at once.  Worker threads build each job's host
 * topology (through the topology cache, gm_set_topology_cache); a scheduler
 * thread admits built jobs in submission order while their device footprint
 * fits the memory budget and a slot is free, and runs rounds: every running
 * job advances one pass per round, the jobs of one (smoother, FP64 pass)
 * group in fused launches, groups on their own streams.  The workers upload
 * and start admitted jobs and copy out finished ones while rounds run;
 * started jobs join the next round (continuous batching).  A finished job's
 * device memory is released at once.  Each job's result is
 * bit-identical to gm_mesh_smooth on that mesh alone (same smoother, passes
 * and precision, GM_KERNEL_WAVE, spawn coefficient 1/20).  The reference has
 * no serving layer; a job is one AbstractSmoother::smoothMesh call
 * (Smoothers/AbstractSmoother.cpp:57-211).
 *
 * budget_bytes <= 0: all device memory free at creation less 1 GiB (a larger
 * request is clamped to that).  A job whose footprint exceeds the budget fails
 * with GM_ERR_NOMEM.  Crew: evaluator, sampler and scaling for every job; the
 * texture sampler's grid comes from gm_server_set_texture before the first
 * job.  Destroying the server cancels unfinished jobs. */
typedef struct gm_server gm_server;

int gm_server_create(int64_t budget_bytes, int max_active, int worker_threads, int evaluator, int sampler,
                     double scaling, gm_server** out);
int gm_server_destroy(gm_server* server);
/* The texture grid every job of a GM_TEXTURE server samples (arguments as
 * gm_mesh_set_crew); only before the first job is submitted.  Replaces
 * updateCudaSamplerTextures (Samplers/TextureSampler.cpp:75-80) for the
 * served meshes. */
int gm_server_set_texture(gm_server* server, const int32_t* tsize, const double* tmin, const double* text,
                          const float* texels);
/* Launch shape for `smoother` (as gm_batch_set_shape). */
int gm_server_set_shape(gm_server* server, int smoother, int patches_per_group, int tile_elems);
/* Queues a job for the mesh (caller numbering); *job = its id.  borrow = 0
 * copies the arrays before returning; borrow != 0 reads them in place, so
 * the caller keeps them alive and unchanged until gm_server_wait reports the
 * job built, done or failed. */
int gm_server_submit(gm_server* server, int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets,
                     const uint32_t* tets, int64_t nhexs, const uint32_t* hexs, int smoother, int passes,
                     int precision, int borrow, int64_t* job);
/* Queues a job encoded with gm_job_encode (copied before returning). */
int gm_server_submit_job(gm_server* server, const void* job_bytes, int64_t bytes, int64_t* job);
/* Waits until the job finished or timeout_ms passed (< 0: no limit);
 * *state = 0 queued, 1 built, 2 active, 3 done, 4 failed. */
int gm_server_wait(gm_server* server, int64_t job, double timeout_ms, int* state);
/* Results of a finished job: pos[3 nverts] in the caller's numbering,
 * stats[3 (passes+1)] and hist[20 (passes+1)] as gm_mesh_smooth, times4 =
 * {submit->built, built->admitted, admitted->done, submit->done} in ms (any
 * may be NULL).  A failed job returns its error. */
int gm_server_result(gm_server* server, int64_t job, double* pos, double* stats, int32_t* hist, double* times4);
/* Forgets a finished job. */
int gm_server_release(gm_server* server, int64_t job);
/* out8 = {submitted, done, failed, rounds, peak active jobs, peak reserved
 * bytes, budget bytes, wall ms of the rounds}. */
int gm_server_stats(gm_server* server, double* out8);

/* ---- Job files: a lossless binary form of a smoothing request (mesh in
 * the caller's numbering, smoother, passes, precision), for queues and the
 * serving loop.  The reference's JSON meshes (Serialization/JsonSerializer.cpp:25-120,
 * JsonDeserializer.cpp:38-115) print positions with 6 significant digits
 * and keep a fixed vertex only as a constraint id; a job keeps every bit.
 * Layout: "GMJ1", a 64-byte header (counts, request), positions f64,
 * fixed flags u8, tetrahedra u32, hexahedra u32 (8-byte aligned sections),
 * then an FNV-1a 64 checksum of everything before it; little-endian.
 * Host only. */
/* Bytes of a job with these counts (-1 for bad counts). */
int64_t gm_job_size(int64_t nverts, int64_t ntets, int64_t nhexs);
int gm_job_encode(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets, const uint32_t* tets,
                  int64_t nhexs, const uint32_t* hexs, int smoother, int passes, int precision, void* out,
                  int64_t out_bytes);
/* Checks magic, sizes and checksum; counts3 = {nverts, ntets, nhexs},
 * request3 = {smoother, passes, precision} (either may be NULL). */
int gm_job_info(const void* job_bytes, int64_t bytes, int64_t* counts3, int32_t* request3);
/* pos[3 nverts], fixed[nverts] (0/1), tets[4 ntets], hexs[8 nhexs]. */
int gm_job_decode(const void* job_bytes, int64_t bytes, double* pos, int32_t* fixed, uint32_t* tets,
                  uint32_t* hexs);

/* Synthetic input: an n^3-cell unit-cube mesh, kind 0 = Kuhn tetrahedra
 * (6 per cell, det(v3-v0, v3-v1, v3-v2) > 0), 1 = hexahedra (v0..v3 bottom
 * counter-clockwise, v4..v7 top, Meshers/DebugMesher.cpp:285-293); vertices
 * numbered x-fastest (DebugMesher.cpp:173-181), boundary vertices fixed,
 * interior vertices jittered per axis by U[-jitter/n, jitter/n] from
 * SplitMix64(seed, 3 v + axis), with the jitter of an element's vertices
 * halved until every orientation determinant is at least min_det_fraction
 * of the unjittered one.  Host only.  pos[3 (n+1)^3], fixed[(n+1)^3],
 * elems[4 * 6 n^3] (tets) or [8 n^3] (hexes). */
int gm_cube_mesh(int n, int kind, double jitter, uint64_t seed, double min_det_fraction, double* pos,
                 int32_t* fixed, uint32_t* elems);

/* ---- Allocation-free evaluation.
 *
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It computes what gm_mesh_evaluate
 * computes (the FP64 min, sum of 1/q and 20-bucket histogram of the element
 * qualities; same bits) over arrays the caller already holds on the current
 * device: d_pos[3 nverts] FP64 positions, d_tets[4 ntets], d_hexs[8 nhexs]
 * (any vertex numbering), with the mean-ratio or metric-conformity evaluator
 * and the uniform or analytic sampler at `scaling`.  Results go to device
 * memory: d_out3 = {min, sum of 1/q (+inf if any q <= 0), element count},
 * d_hist20[20].  The kernels are queued on `stream` (a hipStream_t; NULL =
 * the default stream) and the call returns without waiting.  It calls no
 * hipMalloc, hipFree or allocating routine: the reduction's partials live in
 * a static device array (about 57 KB, held by each GPU the library runs on
 * until the process ends), so calls that overlap on one device race (queue
 * them on one stream).  Sizes beyond GM_NOALLOC_MAX_VERTS vertices or
 * GM_NOALLOC_MAX_ELEMS elements are refused with GM_ERR_ARG. */
#define GM_NOALLOC_MAX_VERTS 4294967295LL
#define GM_NOALLOC_MAX_ELEMS 2147483648LL
int gm_evaluate_noalloc(const double* d_pos, int64_t nverts, const uint32_t* d_tets, int64_t ntets,
                        const uint32_t* d_hexs, int64_t nhexs, int evaluator, int sampler, double scaling,
                        double* d_out3, int32_t* d_hist20, void* stream);

/* Device math probe (for tests): op 0 sqrt(a), 1 a/b, 2 a^(2/3), 3 a^(1/3),
 * 4 exp(a), 5 a^(-2/3) as the FP32-polish and evaluation-record kernels
 * compute it (hardware seed, one Halley step), 6 cos(a), 7 ln(a), 8 a^b,
 * all in FP64 on the device; n values. */
int gm_math_probe(int op, const double* a, const double* b, double* out, int64_t n);

/* Issue-rate probe of the current device's FP32 vector instructions (the
 * ceiling the affine smoothing kernels are measured against): op 0 chains of
 * the transcendental pair exp2/log2 of a^(-2/3), op 1 chains of FMAs, op 2
 * both on independent chains in one loop.  rates = {transcendental lane ops
 * per second, FMA lane ops per second, best kernel ms}.  No counterpart in
 * the reference. */
int gm_throughput_probe(int op, double* rates);

#ifdef __cplusplus
}
#endif

#endif /* GPUMESH_H */
