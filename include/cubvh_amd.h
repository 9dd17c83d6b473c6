/*
 * cubvh_amd.h -- C ABI of the MI355X (gfx950) BVH query library libcubvh_amd.so.
 *
 * Plain pointers and sizes only.  Every entry point returns CUBVH_OK (0) or a
 * negative status; the message of the last failure on the calling thread is
 * available from cubvh_last_error().  Nothing throws across this boundary and no
 * entry point takes ownership of a caller pointer.
 *
 * Device entry points take device pointers that stay valid until the work queued
 * on `stream` (a hipStream_t; NULL is the legacy default stream) has completed.
 * They enqueue asynchronously on that stream and do not synchronize, except where
 * noted.  `device` is the HIP ordinal the pointers live on.
 *
 * Reference interfaces replaced (ashawkey/cubvh):
 *   src/bindings.cpp:14-26      pybind11 module _cubvh (cuBVH class, create_cuBVH,
 *                               from_state_dict, floodfill)
 *   include/gpu/api_gpu.h:17-32 abstract cubvh::cuBVH and its factories
 *   src/api_gpu.cu:17-152       cuBVHImpl (gather, lazy upload, query methods)
 */
#ifndef CUBVH_AMD_H
#define CUBVH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUBVH_OK 0
#define CUBVH_ERR_ARG (-1)    /* invalid argument (null pointer, bad size, bad index) */
#define CUBVH_ERR_HIP (-2)    /* a HIP runtime call failed */
#define CUBVH_ERR_STATE (-3)  /* malformed serialized state */
#define CUBVH_ERR_NOMEM (-4)  /* host allocation failed */

/* Opaque BVH over one triangle mesh (host copy plus lazily uploaded device copies). */
typedef struct cubvh_bvh cubvh_bvh;

/* Message of the last failure on this thread ("" if none). */
const char* cubvh_last_error(void);

/* ABI version: (major << 16) | minor. */
int cubvh_abi_version(void);

/* ---------------------------------------------------------------------------
 * Construction (host only: no GPU is touched, as in the reference's lazy upload).
 * Replaces create_cuBVH(vertices, triangles)   src/api_gpu.cu:128-130,
 *          cuBVHImpl(vertices, triangles)      src/api_gpu.cu:21-44,
 *          TriangleBvh4::build                 src/bvh.cu:515-620.
 * vertices: fp32 [n_vertices, 3] row-major; faces: int32 [n_faces, 3] row-major.
 * Face indices are bounds-checked (the reference does not check them).
 * leaf_size: bits 0-7 = maximum triangles per leaf, 1..16 (0 selects the default, 16: distance
 * queries run fastest on it; leaves of 8 suit ray tracing better);
 * bit 8 set = split at the object median instead of the binned surface-area
 * heuristic (the default).  n_faces <= 2^27 - 2 (a leaf reference holds a 27-bit
 * first-face index and a 4-bit count); a forest's meshes together likewise.
 * ------------------------------------------------------------------------- */
int cubvh_bvh_create(const float* vertices, int64_t n_vertices, const int32_t* faces, int64_t n_faces,
                     int32_t leaf_size, cubvh_bvh** out);

/* Construction on the GPU (no counterpart in the reference, whose build is the
 * host-only TriangleBvh4::build, src/bvh.cu:515-620): vertices fp32 [n_vertices, 3]
 * and faces int32 [n_faces, 3] in the memory of `device`.  leaf_size: bits 0-7 = faces
 * per leaf, 1..16 (0 = 16); bits 8 and 9 clear = object-median splits (every run of faces
 * halved at the median centroid of its longest centroid axis, level by level: the
 * leaves of the host build with bit 8 set); bit 9 set = surface-area heuristic (every
 * run split at the slot of least cost over its three centroid-sorted face lists: the
 * slowest GPU build, the fastest tree to query; a tree too deep for the walk is
 * rebuilt with median splits); bit 8 set = linear BVH from 63-bit Morton codes
 * (fastest to build, slowest to query); both bits set is an error.  All collapse to
 * the same 4-wide layout.  The
 * tree stays on `device`; nodes and triangles are copied to the host only when a
 * host-side use needs them (export, a forest, another device).  Queries on it return the same results as on a
 * host-built tree.  Face indices and coordinates are checked, and so is the built
 * structure (every face in exactly one leaf).  Synchronizes the stream. */
int cubvh_bvh_create_gpu(int device, const float* vertices, int64_t n_vertices, const int32_t* faces,
                         int64_t n_faces, int32_t leaf_size, cubvh_bvh** out, void* stream);

/* Frees the host and every device copy.  Synchronizes the devices that hold one. */
void cubvh_bvh_destroy(cubvh_bvh* bvh);

/* Sizes of the serialized form and of the tree:
 * info[0] = n_nodes (128-byte 4-wide nodes), info[1] = n_faces, info[2] = tree depth
 * (root = 1), info[3] = stack entries the traversals need, info[4] = leaf size. */
int cubvh_bvh_info(const cubvh_bvh* bvh, int64_t info[5]);

/* ---------------------------------------------------------------------------
 * Serialization.  Replaces cuBVH::state_dict / from_state_dict
 * (include/gpu/bvh.cuh:84-109, src/api_gpu.cu:47-68,110-121).
 * Own layout: nodes int32 [n_nodes, 32] (one 128-byte node per row) and
 * triangles int32 [n_faces, 12] (the 48-byte record: a, b, c as fp32, 4 zero bytes,
 * int64 original face index -- the same record layout as the reference).
 * ------------------------------------------------------------------------- */
int cubvh_bvh_export(const cubvh_bvh* bvh, int32_t* nodes, int32_t* triangles);
int cubvh_bvh_import(const int32_t* nodes, int64_t n_nodes, const int32_t* triangles, int64_t n_faces,
                     cubvh_bvh** out);

/* Exports in the reference's state_dict layout (the inverse of its from_state_dict,
 * src/api_gpu.cu:47-68; nodes bvh.cuh:14-21, escape links src/bvh.cu:595-614):
 * nodes9 int32 [*n_nodes9, 9], triangles int32 [n_faces, 12] (leaf order).  Every
 * wide node becomes 4 consecutive reference nodes (empty slots: empty leaves with an
 * empty box), so the reference's traversals walk this tree.  With nodes9 NULL, only
 * writes the node count needed to *n_nodes9 (1 + 4 * n_nodes). */
int cubvh_bvh_export_reference(const cubvh_bvh* bvh, int32_t* nodes9, int64_t* n_nodes9, int32_t* triangles);

/* Imports a state_dict written by the reference: nodes int32 [n_nodes, 9]
 * (TriangleBvhNode, bvh.cuh:14-21) and triangles int32 [n_faces, 12] (Triangle,
 * triangle.cuh:174-175).  The triangles (with their original face ids) are rebuilt
 * into this library's layout; the reference nodes are validated and then dropped. */
int cubvh_bvh_import_reference(const int32_t* nodes9, int64_t n_nodes, const int32_t* triangles12,
                               int64_t n_faces, int32_t leaf_size, cubvh_bvh** out);

/* Mesh in original face order, recovered from the triangle records:
 * tri_vertices fp32 [n_faces, 3, 3]. */
int cubvh_bvh_triangles(const cubvh_bvh* bvh, float* tri_vertices);

/* Copies the BVH to `device` now (otherwise done on the first query, as the
 * reference's lazy upload at src/api_gpu.cu:76-78).  Synchronous. */
int cubvh_bvh_upload(cubvh_bvh* bvh, int device, void* stream);

/* ---------------------------------------------------------------------------
 * Queries.  n is the number of queries.  face_id is the ORIGINAL face index
 * (int64).  Ties are broken toward the lowest face index.
 * ------------------------------------------------------------------------- */

/* Replaces cuBVH::unsigned_distance (src/api_gpu.cu:83-95, src/bvh.cu:468-488,
 * unsigned_distance_kernel src/bvh.cu:687-712).
 * positions fp32 [n,3]; distances fp32 [n]; face_id int64 [n]; uvw fp32 [n,3] or NULL.
 * A query with no face within 1000 gets distance 0 and face 0. */
int cubvh_unsigned_distance(cubvh_bvh* bvh, int device, int64_t n, const float* positions, float* distances,
                            int64_t* face_id, float* uvw, void* stream);

/* No counterpart in the reference: cubvh_unsigned_distance for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results, bit for bit, as
 * cubvh_unsigned_distance; at most 2^20 queries per call (larger n is refused).  The
 * Morton order of the queries and its workspace live in a static 32 MiB device array
 * that each GPU running the library holds until the process ends; the call allocates
 * and frees nothing.  The tree must already be on `device` (cubvh_bvh_upload or an
 * earlier query there), otherwise the call fails.  Every call on one GPU uses that
 * GPU's one array: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialized by the caller. */
int cubvh_unsigned_distance_noalloc(cubvh_bvh* bvh, int device, int64_t n, const float* positions,
                                    float* distances, int64_t* face_id, float* uvw, void* stream);

/* Replaces cuBVH::signed_distance (src/api_gpu.cu:97-108, src/bvh.cu:430-466,
 * kernels src/bvh.cu:631-685).  mode 0 = watertight, any other value = raystab.
 * The raystab random offsets are keyed by the query index within this call. */
int cubvh_signed_distance(cubvh_bvh* bvh, int device, int64_t n, uint32_t mode, const float* positions,
                          float* distances, int64_t* face_id, float* uvw, void* stream);

/* The two steps of cubvh_unsigned_distance / watertight cubvh_signed_distance for
 * callers that query one point set repeatedly (no counterpart in the reference,
 * whose kernels take queries in input order).  cubvh_morton_order writes a
 * permutation of 0..n-1 (uint32 [n], n < 2^31) that puts the points in Morton order
 * of their bounding box; cubvh_distance_ordered walks the queries in the given
 * order as wave packets.  kind 0 = unsigned, 1 = signed (watertight).  Results are
 * identical to the unordered calls for any permutation; order entries >= n are
 * skipped (their queries are not written).  Neither call keeps a reference to the
 * caller's buffers; cubvh_morton_order allocates stream-ordered scratch,
 * cubvh_distance_ordered allocates nothing. */
int cubvh_morton_order(int device, int64_t n, const float* positions, uint32_t* order, void* stream);
int cubvh_distance_ordered(cubvh_bvh* bvh, int device, int64_t n, int32_t kind, const float* positions,
                           const uint32_t* order, float* distances, int64_t* face_id, float* uvw, void* stream);

/* Replaces cuBVH::ray_trace (src/api_gpu.cu:70-81, src/bvh.cu:490-513,
 * raytrace_kernel src/bvh.cu:714-740).  rays_o, rays_d, positions fp32 [n,3];
 * face_id int64 [n] (-1 on a miss); depth fp32 [n] (1000 on a miss). */
int cubvh_ray_trace(cubvh_bvh* bvh, int device, int64_t n, const float* rays_o, const float* rays_d,
                    float* positions, int64_t* face_id, float* depth, void* stream);

/* Replaces floodfill (src/api_gpu.cu:136-152, _floodfill_batch
 * include/gpu/floodfill.cuh:157-272).  grid: uint8/bool [B,H,W,D] (nonzero =
 * occupied); labels int32 [B,H,W,D]: -1 for occupied voxels, otherwise the minimum
 * linear index (global across the batch) of the voxel's 6-connected free component,
 * in tensor order (last axis fastest).  B*H*W*D must be < 2^31.  Works in `labels`
 * only (no scratch memory). */
int cubvh_floodfill(int device, const uint8_t* grid, int32_t B, int32_t H, int32_t W, int32_t D,
                    int32_t* labels, void* stream);

/* No counterpart in the reference: backward of the unsigned distance.
 * For query i with closest point c_i = uvw_i . (v_a, v_b, v_c) of face face_id[i]
 * and n_i = (p_i - c_i) / |p_i - c_i| (zero when the distance is 0):
 *   grad_points[i]   = grad_dist[i] * n_i
 *   grad_vertices[v] += -grad_dist[i] * uvw_i[k] * n_i  for the k-th vertex v of the face
 * vertices fp32 [n_vertices,3], faces int32 [n_faces,3] (original mesh, on device);
 * grad_points fp32 [n,3] may be NULL; grad_vertices fp32 [n_vertices,3] may be NULL and
 * is accumulated into (the caller zeroes it).  The vertex sums are gathered in a fixed
 * order without atomics (per face its queries by ascending index, summed in chunks of 32
 * whose sums are reduced in runs of 32, each added in order, until one is left; per
 * vertex its corners 3 f + k ascending): the same bits on every call.  With grad_vertices the call
 * allocates stream-ordered scratch (sorts of the n face ids and 3 n_faces corners);
 * n < 2^31 and 3 n_faces < 2^31. */
int cubvh_udf_backward(int device, int64_t n, const float* grad_dist, const float* positions,
                       const int64_t* face_id, const float* uvw, const float* vertices, int64_t n_vertices,
                       const int32_t* faces, int64_t n_faces, float* grad_points, float* grad_vertices,
                       void* stream);

/* No counterpart in the reference: the vertex -> corners table of a mesh, which the
 * vertex gradients read and which depends only on the faces.  corners uint32 [3 n_faces]
 * = the corners 3 f + k grouped by vertex, ascending inside a vertex (corners with an
 * out-of-range vertex index last); vertex v owns corners[corner_begin[v] ..
 * corner_end[v]) (uint32 [n_vertices] each; 0, 0 for an unused vertex).  Device buffers
 * on `device`; allocates stream-ordered scratch for the sort.  3 n_faces < 2^31. */
int cubvh_vertex_corners(int device, const int32_t* faces, int64_t n_faces, int64_t n_vertices, uint32_t* corners,
                         uint32_t* corner_begin, uint32_t* corner_end, void* stream);

/* cubvh_udf_backward with the mesh's vertex -> corners table built beforehand by
 * cubvh_vertex_corners from the same faces (for callers that run many backward passes
 * over one mesh: each call then sorts only its n face ids).  The same bits as
 * cubvh_udf_backward.  The table is device data and is not checked beyond this: table
 * entries outside the 3 n_faces corners are skipped, so a table of other faces gives
 * wrong vertex gradients but reads no memory outside it. */
int cubvh_udf_backward_with_corners(int device, int64_t n, const float* grad_dist, const float* positions,
                                    const int64_t* face_id, const float* uvw, const float* vertices,
                                    int64_t n_vertices, const int32_t* faces, int64_t n_faces, float* grad_points,
                                    float* grad_vertices, const uint32_t* corners, const uint32_t* corner_begin,
                                    const uint32_t* corner_end, void* stream);

/* ---------------------------------------------------------------------------
 * Batched queries over many meshes.  No counterpart in the reference (which takes
 * one mesh per cuBVH and one launch per query call); for callers with many meshes
 * and uneven query sets per mesh.
 * ------------------------------------------------------------------------- */
typedef struct cubvh_forest cubvh_forest;

/* Concatenates the trees of n_meshes BVHs (the BVHs stay owned by the caller and may
 * be destroyed afterwards).  Face ids reported by the forest are the per-mesh
 * original face indices.  Host-built trees are concatenated on the host and uploaded
 * on first use.  When every tree was built by cubvh_bvh_create_gpu on one device and
 * none has been copied to the host, the forest is assembled on that device (device
 * copies, synchronous) and serves queries on that device only. */
int cubvh_forest_create(cubvh_bvh* const* bvhs, int32_t n_meshes, cubvh_forest** out);
void cubvh_forest_destroy(cubvh_forest* forest);

/* info[0] = n_meshes, info[1] = total nodes, info[2] = total faces, info[3] = stack
 * entries needed. */
int cubvh_forest_info(const cubvh_forest* forest, int64_t info[4]);

/* Uploads the forest to `device` and allocates the query workspace for up to
 * max_queries queries and Morton-prefix bins of up to max_order_bits bits
 * (n_meshes * 2^max_order_bits <= 2^30).  After
 * this call, cubvh_forest_unsigned_distance within those limits allocates nothing
 * and issues only kernel launches (safe to capture in a hipGraph).  Synchronous. */
int cubvh_forest_reserve(cubvh_forest* forest, int device, int64_t max_queries, int32_t max_order_bits,
                         void* stream);

/* Unsigned distance of n queries, positions fp32 [n,3] concatenated mesh by mesh;
 * query_offsets: int64 [n_meshes + 1] ON THE DEVICE, queries of mesh m are
 * [query_offsets[m], query_offsets[m+1]).  The offsets must be a CSR of the n queries
 * (offsets[0] = 0, non-decreasing, offsets[n_meshes] = n); they are read on the device
 * only, so other offsets are not reported: they give wrong answers, but every write
 * stays inside the outputs and the workspace.  Outputs as cubvh_unsigned_distance.
 * tuning: n_tuning (0..5) entries, the missing ones default to {12, 1, 8, 4, 0}:
 * [0] Morton-prefix bits used to order the queries inside each mesh (0..16, 0 keeps
 * input order; n_meshes * 2^bits <= 2^30), [1] 1 = persistent grid taking 64-query
 * tiles from an atomic counter, 0 = one wave per tile, [2] waves per CU of the
 * persistent grid (1..32), [3] waves per workgroup (1, 2, 4 or 8), [4] for the
 * one-wave-per-tile grid, runs of that many consecutive workgroups given to each XCD
 * (0..65536, 0 = dispatch order; neighbouring tiles then share one L2).  Grows the workspace
 * when the reserved one is too small, which is refused while the stream is being
 * captured. */
int cubvh_forest_unsigned_distance(cubvh_forest* forest, int device, int64_t n, const float* positions,
                                   const int64_t* query_offsets, float* distances, int64_t* face_id, float* uvw,
                                   const int32_t* tuning, int32_t n_tuning, void* stream);

/* ---------------------------------------------------------------------------
 * Sparse-voxel hash table: a static open-addressed table of int32 keys of D words
 * (1 <= D <= 63), linear probing.  Replaces the reference's cuHashTable
 * (include/gpu/api_gpu.h:37-52, src/api_gpu.cu:215-281, include/gpu/hashtable.cuh,
 * bound at src/bindings.cpp:34-44).
 * Differences, all on purpose: the table stores the keys (the reference keeps a
 * pointer to the caller's coords, so its table breaks when they are freed or when a
 * second insert() replaces them); duplicate keys answer with their lowest row index
 * (the reference's GPU table answers with any one of them, its CPU table with the
 * lowest); capacities and key widths are checked.
 * ------------------------------------------------------------------------- */
typedef struct cubvh_hashtable cubvh_hashtable;

/* Empty table for keys of num_dims words (capacity 0: nothing allocated).
 * Replaces create_cuHashTable() + set_num_dims(d)   src/api_gpu.cu:278-280. */
int cubvh_hashtable_create(int32_t num_dims, cubvh_hashtable** out);
/* Frees the table's device memory (synchronizes its device). */
void cubvh_hashtable_destroy(cubvh_hashtable* ht);

/* Key width expected by the next insert.   Replaces set_num_dims   api_gpu.cu:222-224. */
int cubvh_hashtable_set_num_dims(cubvh_hashtable* ht, int32_t num_dims);
/* info[0] = num_dims (get_num_dims, api_gpu.cu:226-228), info[1] = capacity (slots),
 * info[2] = key words the table memory holds, info[3] = int32 words per slot,
 * info[4] = device of the table memory (-1 if none). */
int cubvh_hashtable_info(cubvh_hashtable* ht, int64_t info[5]);

/* capacity slots on `device`, all empty (0 frees the table).  Reallocates only when
 * the current memory is too small or on another device (hipMalloc / hipFree).
 * Replaces resize(capacity)   api_gpu.cu:230-232 (whose slots stay uninitialised). */
int cubvh_hashtable_resize(cubvh_hashtable* ht, int device, int64_t capacity, void* stream);
/* Marks every slot empty (laid out for the current num_dims); no-op at capacity 0.
 * Replaces prepare()   api_gpu.cu:234-237, hashtable.cuh:124-131,181-186. */
int cubvh_hashtable_prepare(cubvh_hashtable* ht, void* stream);
/* Inserts the n rows of coords int32 [n, num_dims] (device memory on `device`), row i
 * with index i; sets num_dims.  No-op at capacity 0 or n == 0.  A row that finds no
 * free slot (table full) is dropped, as in the reference.  Keys of another width than
 * the ones already stored are refused until prepare().
 * Replaces insert(coords)   api_gpu.cu:239-249, hashtable.cuh:136-158,188-195. */
int cubvh_hashtable_insert(cubvh_hashtable* ht, int device, const int32_t* coords, int64_t n, int32_t num_dims,
                           void* stream);
/* capacity = max(16, 2 n), clear, insert (n < 2^30).
 * Replaces build(coords)   api_gpu.cu:251-261, hashtable.cuh:197-206. */
int cubvh_hashtable_build(cubvh_hashtable* ht, int device, const int32_t* coords, int64_t n, int32_t num_dims,
                          void* stream);
/* out int32 [m]: the lowest row index stored with key queries[i] (int32 [m, num_dims],
 * num_dims equal to the stored keys' width), or -1.  Fails on a table never built.
 * Replaces search(queries)   api_gpu.cu:263-274, hashtable.cuh:161-176,208-214. */
int cubvh_hashtable_search(cubvh_hashtable* ht, int device, const int32_t* queries, int64_t m, int32_t num_dims,
                           int32_t* out, void* stream);

/* ---------------------------------------------------------------------------
 * Sparse marching cubes.  Replaces sparse_marching_cubes(coords, corners, iso,
 * ensure_consistency)   src/api_gpu.cu:154-211, include/gpu/spcumc.cuh:631-868
 * (bound at src/bindings.cpp:28-31), in two calls so that the caller allocates the
 * outputs once their sizes are known.
 * Output (the reference's): one vertex per lattice edge with a sign change, in
 * ascending (x, y, z, axis) order of the edge's lowest lattice point, placed by linear
 * interpolation from the first cell (input order) holding the edge; per cell in input
 * order, the triangles of its case (corner value < iso = inside) with the second and
 * third vertex exchanged.  ensure_consistency replaces each corner value by the mean
 * over the cells sharing the lattice point, summed in double in cell order (the
 * reference's CPU path; its GPU path sums in fp32 in an unspecified order).
 * ------------------------------------------------------------------------- */
typedef struct cubvh_smc cubvh_smc;

/* coords int32 [n, 3] (lowest lattice point of each cell), corners fp32 [n, 8] (corner
 * k at offset ((k ^ k >> 1) & 1, k >> 1 & 1, k >> 2)), both on `device`, valid until
 * cubvh_smc_write has been enqueued.  n <= 2^28; the cells may span at most 2^20 - 2
 * lattice steps per axis.  Runs the extraction up to the vertex and triangle counts
 * (synchronizes the stream) and returns a handle holding its scratch. */
int cubvh_sparse_marching_cubes(int device, const int32_t* coords, const float* corners, int64_t n, float iso,
                                int32_t ensure_consistency, cubvh_smc** out, int64_t* n_vertices,
                                int64_t* n_triangles, void* stream);
/* Writes vertices fp32 [n_vertices, 3] and triangles int32 [n_triangles, 3] (device
 * buffers on the extraction's device) on the extraction's stream; once per handle. */
int cubvh_smc_write(cubvh_smc* smc, float* vertices, int32_t* triangles);
/* Releases the scratch (stream-ordered, after the writes) and the handle. */
void cubvh_smc_destroy(cubvh_smc* smc);

/* ---------------------------------------------------------------------------
 * Measurement aid (no counterpart in the reference): the device's VALU issue
 * ceiling.  Runs 8 waves on every SIMD, each lane 16 independent fp32 fma chains of
 * `iters` steps (1 <= iters <= 2^20), on the null stream, and synchronizes.  form 0:
 * operands in VGPRs and literals (the fastest fp32 form); form 1: one operand in an
 * SGPR (the form of arithmetic on wave-uniform node and face records), which issues
 * at about half the rate on gfx950.  *wave_instructions = the fma instructions issued
 * (counted per wave64), *ms = the kernel time.  The walk kernels are bounded by VALU
 * issue; their SQ_INSTS_VALU per second over these rates places them between the two.
 * ------------------------------------------------------------------------- */
int cubvh_valu_probe(int device, int32_t form, int32_t iters, double* wave_instructions, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* CUBVH_AMD_H */
