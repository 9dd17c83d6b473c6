// capi.cpp -- the C ABI declared in include/cubvh_amd.h.
//
// Host side of the library: argument checks, the host BVH (built without touching
// any GPU), lazy per-device upload, and kernel launches on the caller's stream.
// Every entry point catches C++ exceptions and reports a status code instead.
#include "cubvh_amd.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "bvh.hpp"
#include "forest.hpp"
#include "gpubuild.hpp"
#include "hashtable.hpp"
#include "smc.hpp"
#include "query.hpp"

using namespace cubvh_amd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
int fail_hip(hipError_t e, const char* what) {
    return fail(CUBVH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int kMaxDevices = 64;
constexpr int kDefaultLeaf = 16;  // faces per leaf unless the caller asks otherwise
// Point-query batches at least this large are sorted into Morton order and walked as
// wave packets; smaller ones are walked one lane per query in input order.
constexpr int64_t kMortonMin = 1 << 11;

// Keeps freed stream-ordered allocations in the device's default pool instead of
// returning them to the driver at every synchronization.
int ensure_pool(int device) {
    static std::mutex mu;
    static bool done[64] = {false};
    std::lock_guard<std::mutex> lock(mu);
    if (done[device]) return CUBVH_OK;
    hipMemPool_t pool;
    hipError_t e = hipDeviceGetDefaultMemPool(&pool, device);
    if (e != hipSuccess) return fail_hip(e, "hipDeviceGetDefaultMemPool");
    uint64_t thr = UINT64_MAX;
    e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    if (e != hipSuccess) return fail_hip(e, "hipMemPoolSetAttribute");
    done[device] = true;
    return CUBVH_OK;
}

// Makes `device` current for the scope and restores the caller's device.
struct DeviceScope {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit DeviceScope(int device) {
        err = hipGetDevice(&prev);
        if (err == hipSuccess && prev != device) err = hipSetDevice(device);
    }
    ~DeviceScope() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

// Every array of a tree (or of a forest's concatenated trees) in one device
// allocation, each part 256-byte aligned.
struct DeviceTree {
    void* block = nullptr;
    TreeArrays view;

    // Byte sizes of the six parts of a tree with n_nodes nodes and n_faces faces.
    static void sizes(int64_t n_nodes, int64_t n_faces, size_t bytes[6]) {
        bytes[0] = (size_t)n_nodes * sizeof(WideNode);
        bytes[1] = 12 * (size_t)n_faces * sizeof(float);
        bytes[2] = 8 * (size_t)n_faces * sizeof(float);
        bytes[3] = 8 * (size_t)n_faces * sizeof(float);
        bytes[4] = 32 * (size_t)n_nodes * sizeof(float);
        bytes[5] = 36 * (size_t)n_faces * sizeof(float);
    }
    static size_t offsets(const size_t bytes[6], size_t off[6]) {
        size_t total = 0;
        for (int i = 0; i < 6; ++i) {
            off[i] = total;
            total += (bytes[i] + 255) & ~(size_t)255;
        }
        return total;
    }

    // Allocates the block for a tree of the given sizes and sets the view (no copy).
    hipError_t reserve(int64_t n_nodes, int64_t n_faces) {
        size_t bytes[6], off[6];
        sizes(n_nodes, n_faces, bytes);
        const size_t total = offsets(bytes, off);
        hipError_t e = hipMalloc(&block, std::max<size_t>(total, 256));
        if (e != hipSuccess) {
            block = nullptr;
            return e;
        }
        const char* b = (const char*)block;
        view.nodes = (const WideNode*)(b + off[0]);
        view.tris = (const float4*)(b + off[1]);
        view.tbnd = (const float4*)(b + off[2]);
        view.lbnd = (const float4*)(b + off[3]);
        view.cbnd = (const float4*)(b + off[4]);
        view.tpre = (const float4*)(b + off[5]);
        return hipSuccess;
    }

    template <class H>
    hipError_t upload(const H& h, hipStream_t stream) {
        size_t bytes[6], off[6];
        sizes((int64_t)h.nodes.size(), h.n_faces, bytes);
        offsets(bytes, off);
        const void* src[6] = {h.nodes.data(), h.tris.data(), h.tbnd.data(), h.lbnd.data(), h.cbnd.data(), h.tpre.data()};
        hipError_t e = reserve((int64_t)h.nodes.size(), h.n_faces);
        for (int i = 0; i < 6; ++i)
            if (e == hipSuccess && bytes[i])
                e = hipMemcpyAsync((char*)block + off[i], src[i], bytes[i], hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            (void)hipFree(block);
            block = nullptr;
            view = TreeArrays{};
            return e;
        }
        return hipSuccess;
    }
};

struct cubvh_bvh {
    HostBvh host;
    std::mutex mu;
    DeviceTree dev[kMaxDevices];

    ~cubvh_bvh() {
        for (int d = 0; d < kMaxDevices; ++d) {
            if (!dev[d].block) continue;
            DeviceScope g(d);
            (void)hipDeviceSynchronize();
            (void)hipFree(dev[d].block);
        }
    }

    // A tree built on the GPU lives only on that device (mirror_device) until a
    // host-side consumer (state_dict, a forest, another device) needs it: then nodes
    // and triangles are copied back and the derived arrays recomputed on the host.
    int32_t mirror_device = -1;
    int64_t n_nodes = 0;

    int ensure_host_mirror_locked() {
        if (mirror_device < 0) return CUBVH_OK;
        DeviceScope g(mirror_device);
        if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
        const TreeArrays& v = dev[mirror_device].view;
        host.nodes.resize((size_t)n_nodes);
        host.tris.resize(12 * (size_t)host.n_faces);
        hipError_t e = hipMemcpy(host.nodes.data(), v.nodes, host.nodes.size() * sizeof(WideNode), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(host.tris.data(), v.tris, host.tris.size() * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail_hip(e, "copy of the GPU-built tree to the host");
        mirror_device = -1;
        return CUBVH_OK;
    }
    int ensure_host_derived_locked() {
        int rc = ensure_host_mirror_locked();
        if (rc != CUBVH_OK) return rc;
        if (host.tpre.size() != 36 * (size_t)host.n_faces) compute_face_bounds(host);
        return CUBVH_OK;
    }
    int ensure_host_mirror() {
        std::lock_guard<std::mutex> lock(mu);
        return ensure_host_mirror_locked();
    }
    int ensure_host_derived() {
        std::lock_guard<std::mutex> lock(mu);
        return ensure_host_derived_locked();
    }

    // Whether the tree is already on `device` (uploaded or built there).
    bool resident(int device) {
        if (device < 0 || device >= kMaxDevices) return false;
        std::lock_guard<std::mutex> lock(mu);
        return dev[device].block != nullptr;
    }

    // Uploads on first use (the reference's lazy upload, src/api_gpu.cu:76-78).

    int ensure_device(int device, hipStream_t stream) {
        if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
        std::lock_guard<std::mutex> lock(mu);
        if (dev[device].block) return CUBVH_OK;
        DeviceScope g(device);
        if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
        int rc = ensure_host_derived_locked();
        if (rc != CUBVH_OK) return rc;
        hipError_t e = dev[device].upload(host, stream);
        if (e != hipSuccess) return fail_hip(e, "upload");
        return CUBVH_OK;
    }

    QueryArgs args(int device, int64_t n, void* stream) const {
        QueryArgs q;
        q.n = n;
        q.tree = dev[device].view;
        q.face0_slot = host.face0_slot;
        q.stack_need = host.stack_need;
        q.stream = stream;
        return q;
    }
};

struct cubvh_forest {
    std::vector<WideNode> nodes;
    std::vector<float> tris;
    std::vector<float> tbnd;
    std::vector<float> lbnd;
    std::vector<float> cbnd;
    std::vector<float> tpre;
    std::vector<int32_t> mesh_info;  // 4 per mesh: root, face-0 slot, n_faces, 0
    int32_t n_meshes = 0;
    int64_t n_faces = 0;
    int64_t n_nodes = 0;
    int32_t stack_need = 0;
    // >= 0: assembled on that device from trees built there; no host arrays
    int32_t resident_device = -1;
    std::mutex mu;
    struct DevCopy {
        DeviceTree tree;
        int4* mesh_info = nullptr;
        void* ws = nullptr;
        int64_t ws_n = -1;
        int32_t ws_bits = -1;
        int32_t n_cu = 0;
    };
    DevCopy dev[kMaxDevices];

    ~cubvh_forest() {
        for (int d = 0; d < kMaxDevices; ++d) {
            DevCopy& c = dev[d];
            if (!c.tree.block && !c.mesh_info && !c.ws) continue;
            DeviceScope g(d);
            (void)hipDeviceSynchronize();
            (void)hipFree(c.tree.block);
            (void)hipFree(c.mesh_info);
            (void)hipFree(c.ws);
        }
    }

    int upload(int device, hipStream_t stream) {
        DevCopy& c = dev[device];
        if (c.tree.block) return CUBVH_OK;
        if (resident_device >= 0)
            return fail(CUBVH_ERR_ARG, "this forest was assembled on device " + std::to_string(resident_device) +
                                           " from GPU-built trees and is used there only");
        hipError_t e = hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess) return fail_hip(e, "hipDeviceGetAttribute");
        const size_t mb = mesh_info.size() * sizeof(int32_t);
        if (!c.mesh_info) {
            e = hipMalloc(&c.mesh_info, mb);
            if (e == hipSuccess) e = hipMemcpyAsync(c.mesh_info, mesh_info.data(), mb, hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) {
                (void)hipFree(c.mesh_info);
                c.mesh_info = nullptr;
                return fail_hip(e, "forest upload");
            }
        }
        e = c.tree.upload(*this, stream);
        if (e != hipSuccess) return fail_hip(e, "forest upload");
        return CUBVH_OK;
    }

    int reserve(int device, int64_t n, int32_t bits, hipStream_t stream) {
        DevCopy& c = dev[device];
        if (c.ws && c.ws_n >= n && c.ws_bits >= bits) return CUBVH_OK;
        const int64_t nn = std::max<int64_t>(n, c.ws_n);
        const int32_t bb = std::max<int32_t>(bits, c.ws_bits);
        if (c.ws) {
            hipError_t e = hipStreamSynchronize(stream);
            if (e == hipSuccess) e = hipFree(c.ws);
            c.ws = nullptr;
            if (e != hipSuccess) return fail_hip(e, "forest workspace release");
        }
        hipError_t e = hipMalloc(&c.ws, forest_workspace_bytes(nn, n_meshes, bb));
        if (e != hipSuccess) {
            c.ws = nullptr;
            return fail_hip(e, "hipMalloc(forest workspace)");
        }
        c.ws_n = nn;
        c.ws_bits = bb;
        return CUBVH_OK;
    }
};

#define CUBVH_GUARD_BEGIN try {
#define CUBVH_GUARD_END                                                        \
    }                                                                          \
    catch (const std::bad_alloc&) {                                            \
        return fail(CUBVH_ERR_NOMEM, "host allocation failed");               \
    }                                                                          \
    catch (const std::exception& ex) {                                         \
        return fail(CUBVH_ERR_ARG, std::string("unexpected error: ") + ex.what()); \
    }

namespace {

int finish_bvh(cubvh_bvh* b, cubvh_bvh** out) {
    if (b->mirror_device < 0) b->n_nodes = (int64_t)b->host.nodes.size();
    if (stack_capacity_for(b->host.stack_need) < 0) {
        delete b;
        return fail(CUBVH_ERR_STATE, "tree too deep for the traversal stack");
    }
    *out = b;
    g_last_error.clear();
    return CUBVH_OK;
}

// Leaf-order 12-float internal triangle record -> 48-byte exported record.
void export_record(const float* in, int32_t* out) {
    float rec[9] = {in[0], in[1], in[2], in[4], in[5], in[6], in[8], in[9], in[10]};
    int32_t id32;
    std::memcpy(&id32, &in[3], 4);
    int64_t id = id32;
    std::memcpy(out, rec, 36);
    out[9] = 0;
    std::memcpy(out + 10, &id, 8);
}

// kmode: 0 unsigned, 1 watertight, 2 raystab
int run_distance(cubvh_bvh* bvh, int device, int64_t n, int kmode, const float* positions, float* distances,
                 int64_t* face_id, float* uvw, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!bvh) return fail(CUBVH_ERR_ARG, "bvh is null");
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative query count");
    if (n == 0) return CUBVH_OK;
    if (!positions || !distances || !face_id) return fail(CUBVH_ERR_ARG, "null query buffer");
    if (n > ((int64_t)1 << 37)) return fail(CUBVH_ERR_ARG, "too many queries (at most 2^37 per call)");
    int rc = bvh->ensure_device(device, (hipStream_t)stream);
    if (rc != CUBVH_OK) return rc;
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    QueryArgs q = bvh->args(device, n, stream);
    q.in0 = positions;
    q.out_f = distances;
    q.out_id = face_id;
    q.out_v = uvw;
    hipStream_t hs = (hipStream_t)stream;
    // Large batches are walked in Morton order (stream-ordered scratch from the
    // device's default memory pool, released after the launch).
    void* scratch = nullptr;
    if (n >= kMortonMin && n < ((int64_t)1 << 31)) {
        rc = ensure_pool(device);
        if (rc != CUBVH_OK) return rc;
        const size_t ob = ((size_t)n * 4 + 255) & ~(size_t)255;
        // raystab: the stab passes' lists after the order, the ordering workspace after them
        const size_t sb = kmode == 2 ? ((stab_workspace_bytes(n) + 255) & ~(size_t)255) : 0;
        hipError_t e = hipMallocAsync(&scratch, ob + sb + morton_workspace_bytes(n), hs);
        if (e != hipSuccess) return fail_hip(e, "hipMallocAsync(query order)");
        uint32_t* order = (uint32_t*)scratch;
        e = morton_order(n, positions, (char*)scratch + ob + sb, order, hs);
        if (e != hipSuccess) {
            (void)hipFreeAsync(scratch, hs);
            return fail_hip(e, "query ordering");
        }
        q.order = order;
    }
    q.packet = scratch != nullptr;
    hipError_t e;
    if (kmode == 2 && q.packet) {
        // raystab over Morton-ordered queries: packet distance walk, then the stab rays
        e = launch_distance(q, 0);
        if (e == hipSuccess) e = launch_stab_sign(q, (char*)scratch + (((size_t)n * 4 + 255) & ~(size_t)255));
    } else {
        e = launch_distance(q, kmode);
    }
    if (scratch) {
        hipError_t e2 = hipFreeAsync(scratch, hs);
        if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return fail_hip(e, "distance kernel launch");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

}  // namespace

extern "C" {

const char* cubvh_last_error(void) { return g_last_error.c_str(); }

int cubvh_abi_version(void) { return (1 << 16) | 0; }

int cubvh_bvh_create(const float* vertices, int64_t n_vertices, const int32_t* faces, int64_t n_faces,
                     int32_t leaf_size, cubvh_bvh** out) {
    CUBVH_GUARD_BEGIN
    if (!out) return fail(CUBVH_ERR_ARG, "out is null");
    *out = nullptr;
    if (!vertices || !faces) return fail(CUBVH_ERR_ARG, "vertices and faces must not be null");
    if (n_vertices <= 0 || n_faces <= 0) return fail(CUBVH_ERR_ARG, "mesh must have vertices and faces");
    if (n_faces > (int64_t)0x7fffffff) return fail(CUBVH_ERR_ARG, "too many faces");
    const int split_mode = (leaf_size & 0x100) ? 0 : 1;  // bit 8: object median, else SAH
    if (leaf_size & ~0x1ff) return fail(CUBVH_ERR_ARG, "unknown bits in leaf_size");
    leaf_size &= 0xff;
    if (leaf_size == 0) leaf_size = kDefaultLeaf;
    std::vector<float> tri9(9 * (size_t)n_faces);
    std::vector<int64_t> ids((size_t)n_faces);
    for (int64_t f = 0; f < n_faces; ++f) {
        for (int k = 0; k < 3; ++k) {
            int32_t v = faces[3 * f + k];
            if (v < 0 || v >= n_vertices)
                return fail(CUBVH_ERR_ARG, "face " + std::to_string(f) + " references vertex " + std::to_string(v) +
                                               " outside [0, " + std::to_string(n_vertices) + ")");
            std::memcpy(&tri9[9 * f + 3 * k], vertices + 3 * (int64_t)v, 12);
        }
        ids[f] = f;
    }
    cubvh_bvh* b = new cubvh_bvh;
    std::string err = build_bvh(tri9.data(), ids.data(), n_faces, leaf_size, b->host, split_mode);
    if (err.empty() && split_mode == 1 && stack_capacity_for(b->host.stack_need) < 0)  // SAH tree too deep
        err = build_bvh(tri9.data(), ids.data(), n_faces, leaf_size, b->host, 0);
    if (!err.empty()) {
        delete b;
        return fail(CUBVH_ERR_ARG, err);
    }
    return finish_bvh(b, out);
    CUBVH_GUARD_END
}

int cubvh_bvh_create_gpu(int device, const float* vertices, int64_t n_vertices, const int32_t* faces,
                         int64_t n_faces, int32_t leaf_size, cubvh_bvh** out, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!out) return fail(CUBVH_ERR_ARG, "out is null");
    *out = nullptr;
    if (!vertices || !faces) return fail(CUBVH_ERR_ARG, "vertices and faces must not be null");
    if (n_vertices <= 0 || n_faces <= 0) return fail(CUBVH_ERR_ARG, "mesh must have vertices and faces");
    if (n_faces > MAX_LEAF_FIRST || n_vertices > (int64_t)0x7fffffff)
        return fail(CUBVH_ERR_ARG, "too many faces or vertices for the GPU build");
    if (leaf_size < 0 || (leaf_size & ~0x3ff)) return fail(CUBVH_ERR_ARG, "unknown bits in leaf_size");
    if ((leaf_size & 0x300) == 0x300) return fail(CUBVH_ERR_ARG, "leaf_size bits 8 and 9 exclude each other");
    // bit 8: Morton radix tree, bit 9: surface-area heuristic, neither: object median
    const int topology = (leaf_size & 0x100) ? kTopologyLinear : (leaf_size & 0x200) ? kTopologySah : kTopologyMedian;
    leaf_size &= 0xff;
    if (leaf_size > MAX_LEAF_SIZE) return fail(CUBVH_ERR_ARG, "leaf_size must be in 0..16");
    if (leaf_size == 0) leaf_size = kDefaultLeaf;
    if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    hipStream_t hs = (hipStream_t)stream;
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    if (n_faces <= leaf_size) {
        // one leaf: nothing to build on the device; the host build over the copied-back mesh
        std::vector<float> v(3 * (size_t)n_vertices);
        std::vector<int32_t> fc(3 * (size_t)n_faces);
        hipError_t e = hipMemcpyAsync(v.data(), vertices, v.size() * 4, hipMemcpyDeviceToHost, hs);
        if (e == hipSuccess) e = hipMemcpyAsync(fc.data(), faces, fc.size() * 4, hipMemcpyDeviceToHost, hs);
        if (e == hipSuccess) e = hipStreamSynchronize(hs);
        if (e != hipSuccess) return fail_hip(e, "mesh copy");
        return cubvh_bvh_create(v.data(), n_vertices, fc.data(), n_faces, leaf_size, out);
    }
    int rc = ensure_pool(device);
    if (rc != CUBVH_OK) return rc;
    void* ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, gpu_build_workspace_bytes(n_faces, topology), hs);
    if (e != hipSuccess) return fail_hip(e, "hipMallocAsync(build workspace)");
    struct WsGuard {
        void* p;
        hipStream_t s;
        ~WsGuard() {
            (void)hipFreeAsync(p, s);
            (void)hipStreamSynchronize(s);
        }
    } guard{ws, hs};
    const GpuBuildWorkspace w = carve_gpu_build_workspace(ws, n_faces, topology);
    GpuBuildResult res;
    e = gpu_build_tree(vertices, n_vertices, faces, n_faces, leaf_size, topology, w, res, hs);
    if (e == hipSuccess && topology == kTopologySah && res.too_deep && res.input_errors == 0) {
        // an SAH tree deeper than the stack accepts: the median tree (the host build's
        // fallback too); the SAH workspace holds everything the median build uses
        res = GpuBuildResult();
        e = gpu_build_tree(vertices, n_vertices, faces, n_faces, leaf_size, kTopologyMedian, w, res, hs);
    }
    if (e != hipSuccess) return fail_hip(e, "GPU tree build");
    if (res.input_errors & 1) return fail(CUBVH_ERR_ARG, "a face references a vertex outside [0, n_vertices)");
    if (res.input_errors & 2) return fail(CUBVH_ERR_ARG, "vertex coordinates must be finite");
    if (res.too_deep || stack_capacity_for(3 * res.depth + 1) < 0)
        return fail(CUBVH_ERR_STATE, "GPU-built tree too deep for the traversal stack (depth " +
                                         std::to_string(res.depth) + "); use the host build");
    if (!res.structure_ok) return fail(CUBVH_ERR_STATE, "GPU-built tree failed its structure check");
    cubvh_bvh* b = new cubvh_bvh;
    HostBvh& h = b->host;
    h.n_faces = n_faces;
    h.leaf_size = leaf_size;
    h.depth = res.depth;
    h.stack_need = 3 * res.depth + 1;
    b->n_nodes = res.n_nodes;
    DeviceTree& dt = b->dev[device];
    e = dt.reserve(res.n_nodes, n_faces);
    if (e == hipSuccess) e = gpu_build_arrays(n_faces, w, dt.view, hs);
    if (e == hipSuccess)
        e = hipMemcpyAsync(const_cast<WideNode*>(dt.view.nodes), w.wide, (size_t)res.n_nodes * sizeof(WideNode),
                           hipMemcpyDeviceToDevice, hs);
    if (e == hipSuccess) e = gpu_build_face0_slot(w, h.face0_slot, hs);
    if (e != hipSuccess) {
        delete b;
        return fail_hip(e, "GPU tree arrays");
    }
    b->mirror_device = device;  // nodes and triangles are copied to the host on demand
    return finish_bvh(b, out);
    CUBVH_GUARD_END
}

void cubvh_bvh_destroy(cubvh_bvh* bvh) { delete bvh; }

int cubvh_bvh_info(const cubvh_bvh* bvh, int64_t info[5]) {
    if (!bvh || !info) return fail(CUBVH_ERR_ARG, "null argument");
    info[0] = bvh->n_nodes;
    info[1] = bvh->host.n_faces;
    info[2] = bvh->host.depth;
    info[3] = bvh->host.stack_need;
    info[4] = bvh->host.leaf_size;
    return CUBVH_OK;
}

int cubvh_bvh_export(const cubvh_bvh* bvh, int32_t* nodes, int32_t* triangles) {
    if (!bvh || !nodes || !triangles) return fail(CUBVH_ERR_ARG, "null argument");
    int rc = const_cast<cubvh_bvh*>(bvh)->ensure_host_mirror();  // a cache: logically const
    if (rc != CUBVH_OK) return rc;
    std::memcpy(nodes, bvh->host.nodes.data(), bvh->host.nodes.size() * sizeof(WideNode));
    for (int64_t s = 0; s < bvh->host.n_faces; ++s) export_record(&bvh->host.tris[12 * s], triangles + 12 * s);
    return CUBVH_OK;
}

// The reference's node (include/gpu/bvh.cuh:14-21): box (min, max), left, right,
// escape.  Leaf: left = -begin - 1, right = -end - 1; internal: 4 children at
// [left, right); escape links in pre-order (src/bvh.cu:595-614).
struct RefNodeOut {
    float mn[3], mx[3];
    int32_t left, right, escape;
};
static_assert(sizeof(RefNodeOut) == 36, "reference node is 9 x 4 bytes");

int cubvh_bvh_export_reference(const cubvh_bvh* bvh, int32_t* nodes9, int64_t* n_nodes9, int32_t* triangles) {
    CUBVH_GUARD_BEGIN
    if (!bvh || !n_nodes9) return fail(CUBVH_ERR_ARG, "null argument");
    const int64_t need = 1 + 4 * bvh->n_nodes;  // the root, then 4 children per wide node
    if (!nodes9) {
        *n_nodes9 = need;
        return CUBVH_OK;
    }
    if (*n_nodes9 < need || !triangles) return fail(CUBVH_ERR_ARG, "node buffer too small or triangles null");
    int rc = const_cast<cubvh_bvh*>(bvh)->ensure_host_mirror();  // a cache: logically const
    if (rc != CUBVH_OK) return rc;
    const HostBvh& h = bvh->host;
    std::vector<RefNodeOut> out(1);
    out.reserve((size_t)need);
    float root_mn[3] = {INFINITY, INFINITY, INFINITY}, root_mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    out[0].left = out[0].right = out[0].escape = -1;
    // breadth-first: our wide node w becomes the 4 consecutive reference nodes that
    // follow its parent's entry; empty slots become empty leaves with an empty box
    std::vector<std::pair<int32_t, int64_t>> todo = {{0, 0}};  // (wide node, its reference node)
    for (size_t t = 0; t < todo.size(); ++t) {
        const WideNode& w = h.nodes[(size_t)todo[t].first];
        const int64_t me = todo[t].second;
        const int64_t base = (int64_t)out.size();
        out.resize(out.size() + 4);
        out[(size_t)me].left = (int32_t)base;
        out[(size_t)me].right = (int32_t)(base + 4);
        for (int q = 0; q < 4; ++q) {
            RefNodeOut& r = out[(size_t)(base + q)];
            const int32_t ref = w.child[q];
            r.escape = -1;
            if (ref == CHILD_EMPTY) {
                for (int k = 0; k < 3; ++k) {
                    r.mn[k] = INFINITY;
                    r.mx[k] = -INFINITY;
                }
                r.left = r.right = -1;  // begin == end == 0: no triangles
                continue;
            }
            r.mn[0] = w.lo_x[q]; r.mn[1] = w.lo_y[q]; r.mn[2] = w.lo_z[q];
            r.mx[0] = w.hi_x[q]; r.mx[1] = w.hi_y[q]; r.mx[2] = w.hi_z[q];
            if (ref < 0) {
                const int32_t first = leaf_ref_first(ref);
                const int32_t count = leaf_ref_count(ref);
                r.left = -first - 1;
                r.right = -(first + count) - 1;
            } else {
                todo.push_back({ref, base + q});
            }
            if (me == 0)
                for (int k = 0; k < 3; ++k) {
                    root_mn[k] = std::min(root_mn[k], r.mn[k]);
                    root_mx[k] = std::max(root_mx[k], r.mx[k]);
                }
        }
    }
    for (int k = 0; k < 3; ++k) {
        out[0].mn[k] = root_mn[k];
        out[0].mx[k] = root_mx[k];
    }
    // escape links: pre-order, child order 0..3 (next sibling, else the parent's escape)
    std::vector<std::pair<int64_t, int32_t>> st = {{0, -1}};
    while (!st.empty()) {
        auto [ni, esc] = st.back();
        st.pop_back();
        RefNodeOut& nd = out[(size_t)ni];
        nd.escape = esc;
        if (nd.left < 0) continue;
        for (int c = nd.right - 1; c >= nd.left; --c) st.push_back({c, c + 1 < nd.right ? c + 1 : esc});
    }
    std::memcpy(nodes9, out.data(), out.size() * sizeof(RefNodeOut));
    *n_nodes9 = (int64_t)out.size();
    for (int64_t i = 0; i < h.n_faces; ++i) export_record(&h.tris[12 * i], triangles + 12 * i);
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_bvh_import(const int32_t* nodes, int64_t n_nodes, const int32_t* triangles, int64_t n_faces,
                     cubvh_bvh** out) {
    CUBVH_GUARD_BEGIN
    if (!out) return fail(CUBVH_ERR_ARG, "out is null");
    *out = nullptr;
    if (!nodes || !triangles || n_nodes <= 0 || n_faces <= 0)
        return fail(CUBVH_ERR_STATE, "state has no nodes or no triangles");
    cubvh_bvh* b = new cubvh_bvh;
    HostBvh& h = b->host;
    h.nodes.resize((size_t)n_nodes);
    std::memcpy(h.nodes.data(), nodes, (size_t)n_nodes * sizeof(WideNode));
    h.n_faces = n_faces;
    h.tris.assign(12 * (size_t)n_faces, 0.0f);
    int64_t leaf_max = 1;
    for (int64_t s = 0; s < n_faces; ++s) {
        const int32_t* r = triangles + 12 * s;
        float* o = &h.tris[12 * s];
        float rec[9];
        std::memcpy(rec, r, 36);
        int64_t id;
        std::memcpy(&id, r + 10, 8);
        if (id < 0 || id >= n_faces) {
            delete b;
            return fail(CUBVH_ERR_STATE, "triangle record has an out-of-range face id");
        }
        int32_t id32 = (int32_t)id;
        o[0] = rec[0]; o[1] = rec[1]; o[2] = rec[2];
        std::memcpy(&o[3], &id32, 4);
        o[4] = rec[3]; o[5] = rec[4]; o[6] = rec[5];
        o[8] = rec[6]; o[9] = rec[7]; o[10] = rec[8];
    }
    for (const WideNode& w : h.nodes)
        for (int k = 0; k < 4; ++k)
            if (w.child[k] != CHILD_EMPTY && w.child[k] < 0) leaf_max = std::max<int64_t>(leaf_max, leaf_ref_count(w.child[k]));
    h.leaf_size = (int32_t)leaf_max;
    std::string err = validate_bvh(h);
    if (!err.empty()) {
        delete b;
        return fail(CUBVH_ERR_STATE, "invalid state: " + err);
    }
    compute_face_bounds(h);
    return finish_bvh(b, out);
    CUBVH_GUARD_END
}

int cubvh_bvh_import_reference(const int32_t* nodes9, int64_t n_nodes, const int32_t* triangles12,
                               int64_t n_faces, int32_t leaf_size, cubvh_bvh** out) {
    CUBVH_GUARD_BEGIN
    if (!out) return fail(CUBVH_ERR_ARG, "out is null");
    *out = nullptr;
    if (!nodes9 || !triangles12 || n_nodes <= 0 || n_faces <= 0)
        return fail(CUBVH_ERR_STATE, "state has no nodes or no triangles");
    // Structural check of the reference nodes (bvh.cuh:14-21): internal nodes point to
    // a run of children inside the array, leaves to a triangle range.
    for (int64_t i = 0; i < n_nodes; ++i) {
        int32_t l = nodes9[9 * i + 6], r = nodes9[9 * i + 7];
        if (l < 0) {
            int64_t b = -(int64_t)l - 1, e = -(int64_t)r - 1;
            if (b < 0 || e < b || e > n_faces) return fail(CUBVH_ERR_STATE, "reference leaf range out of bounds");
        } else if (l >= r || r > n_nodes || l == 0) {
            return fail(CUBVH_ERR_STATE, "reference child range out of bounds");
        }
    }
    const int split_mode = (leaf_size & 0x100) ? 0 : 1;
    leaf_size &= 0xff;
    if (leaf_size == 0) leaf_size = kDefaultLeaf;
    std::vector<float> tri9(9 * (size_t)n_faces);
    std::vector<int64_t> ids((size_t)n_faces);
    std::vector<uint8_t> seen((size_t)n_faces, 0);
    for (int64_t s = 0; s < n_faces; ++s) {
        const int32_t* r = triangles12 + 12 * s;
        std::memcpy(&tri9[9 * s], r, 36);
        int64_t id;
        std::memcpy(&id, r + 10, 8);
        if (id < 0 || id >= n_faces || seen[id])
            return fail(CUBVH_ERR_STATE, "reference triangle ids are not a permutation of 0..M-1");
        seen[id] = 1;
        ids[s] = id;
    }
    cubvh_bvh* b = new cubvh_bvh;
    std::string err = build_bvh(tri9.data(), ids.data(), n_faces, leaf_size, b->host, split_mode);
    if (err.empty() && split_mode == 1 && stack_capacity_for(b->host.stack_need) < 0)
        err = build_bvh(tri9.data(), ids.data(), n_faces, leaf_size, b->host, 0);
    if (!err.empty()) {
        delete b;
        return fail(CUBVH_ERR_STATE, err);
    }
    return finish_bvh(b, out);
    CUBVH_GUARD_END
}

int cubvh_bvh_triangles(const cubvh_bvh* bvh, float* tri_vertices) {
    if (!bvh || !tri_vertices) return fail(CUBVH_ERR_ARG, "null argument");
    int rc = const_cast<cubvh_bvh*>(bvh)->ensure_host_mirror();  // a cache: logically const
    if (rc != CUBVH_OK) return rc;
    const HostBvh& h = bvh->host;
    for (int64_t s = 0; s < h.n_faces; ++s) {
        const float* in = &h.tris[12 * s];
        int32_t id;
        std::memcpy(&id, &in[3], 4);
        float* o = tri_vertices + 9 * (int64_t)id;
        o[0] = in[0]; o[1] = in[1]; o[2] = in[2];
        o[3] = in[4]; o[4] = in[5]; o[5] = in[6];
        o[6] = in[8]; o[7] = in[9]; o[8] = in[10];
    }
    return CUBVH_OK;
}

int cubvh_bvh_upload(cubvh_bvh* bvh, int device, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!bvh) return fail(CUBVH_ERR_ARG, "bvh is null");
    return bvh->ensure_device(device, (hipStream_t)stream);
    CUBVH_GUARD_END
}

int cubvh_unsigned_distance(cubvh_bvh* bvh, int device, int64_t n, const float* positions, float* distances,
                            int64_t* face_id, float* uvw, void* stream) {
    return run_distance(bvh, device, n, 0, positions, distances, face_id, uvw, stream);
}

int cubvh_unsigned_distance_noalloc(cubvh_bvh* bvh, int device, int64_t n, const float* positions,
                                    float* distances, int64_t* face_id, float* uvw, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!bvh) return fail(CUBVH_ERR_ARG, "bvh is null");
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative query count");
    if (n == 0) return CUBVH_OK;
    if (!positions || !distances || !face_id) return fail(CUBVH_ERR_ARG, "null query buffer");
    if (n > kNoAllocMaxQueries)
        return fail(CUBVH_ERR_ARG, "at most 2^20 queries per allocation-free call (the size of its static scratch)");
    if (device < 0 || device >= kMaxDevices || device >= kNoAllocDevices)
        return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    if (!bvh->resident(device))
        return fail(CUBVH_ERR_STATE, "the tree is not on this device: call cubvh_bvh_upload first "
                                     "(the allocation-free call does not upload it)");
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    QueryArgs q = bvh->args(device, n, stream);
    q.in0 = positions;
    q.out_f = distances;
    q.out_id = face_id;
    q.out_v = uvw;
    q.packet = false;
    hipStream_t hs = (hipStream_t)stream;
    if (n >= kMortonMin) {
        void* ws = nullptr;
        hipError_t e = noalloc_scratch(device, &ws);
        if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
        const size_t ob = ((size_t)n * 4 + 255) & ~(size_t)255;
        if (ob + morton_workspace_bytes(n) > kNoAllocScratchBytes)
            return fail(CUBVH_ERR_STATE, "the ordering workspace does not fit the static scratch");
        uint32_t* order = (uint32_t*)ws;
        e = morton_order(n, positions, (char*)ws + ob, order, hs);
        if (e != hipSuccess) return fail_hip(e, "query ordering");
        q.order = order;
        q.packet = true;
    }
    hipError_t e = launch_distance(q, 0);
    if (e != hipSuccess) return fail_hip(e, "distance kernel launch");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_valu_probe(int device, int32_t form, int32_t iters, double* wave_instructions, float* ms) {
    CUBVH_GUARD_BEGIN
    if (!wave_instructions || !ms) return fail(CUBVH_ERR_ARG, "null output");
    if (form != 0 && form != 1) return fail(CUBVH_ERR_ARG, "form must be 0 (VGPR operands) or 1 (an SGPR operand)");
    if (iters < 1 || iters > (1 << 20)) return fail(CUBVH_ERR_ARG, "iters must be in [1, 2^20]");
    if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    int n_cu = 0;
    hipError_t e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return fail_hip(e, "hipDeviceGetAttribute");
    const int blocks = n_cu * 8;  // 4 waves per workgroup: 8 waves on each of a CU's 4 SIMDs
    float* out = nullptr;
    e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(float));
    if (e != hipSuccess) return fail_hip(e, "hipMalloc(probe output)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_valu_probe(form, blocks, iters, out, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    hipError_t e2 = hipFree(out);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail_hip(e, "VALU probe");
    *wave_instructions = (double)blocks * 4.0 * (double)iters * (double)kValuProbeChains;
    *ms = t;
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_morton_order(int device, int64_t n, const float* positions, uint32_t* order, void* stream) {
    CUBVH_GUARD_BEGIN
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative query count");
    if (n == 0) return CUBVH_OK;
    if (!positions || !order) return fail(CUBVH_ERR_ARG, "null query buffer");
    if (n >= ((int64_t)1 << 31)) return fail(CUBVH_ERR_ARG, "too many queries to order (at most 2^31 - 1)");
    if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    int rc = ensure_pool(device);
    if (rc != CUBVH_OK) return rc;
    hipStream_t hs = (hipStream_t)stream;
    void* ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, morton_workspace_bytes(n), hs);
    if (e != hipSuccess) return fail_hip(e, "hipMallocAsync(ordering workspace)");
    e = morton_order(n, positions, ws, order, hs);
    hipError_t e2 = hipFreeAsync(ws, hs);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail_hip(e, "query ordering");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_distance_ordered(cubvh_bvh* bvh, int device, int64_t n, int32_t kind, const float* positions,
                           const uint32_t* order, float* distances, int64_t* face_id, float* uvw, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!bvh) return fail(CUBVH_ERR_ARG, "bvh is null");
    if (kind != 0 && kind != 1) return fail(CUBVH_ERR_ARG, "kind must be 0 (unsigned) or 1 (watertight)");
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative query count");
    if (n == 0) return CUBVH_OK;
    if (!positions || !order || !distances || !face_id) return fail(CUBVH_ERR_ARG, "null query buffer");
    if (n >= ((int64_t)1 << 31)) return fail(CUBVH_ERR_ARG, "too many queries (at most 2^31 - 1 per ordered call)");
    int rc = bvh->ensure_device(device, (hipStream_t)stream);
    if (rc != CUBVH_OK) return rc;
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    QueryArgs q = bvh->args(device, n, stream);
    q.in0 = positions;
    q.order = order;
    q.out_f = distances;
    q.out_id = face_id;
    q.out_v = uvw;
    q.packet = true;
    hipError_t e = launch_distance(q, kind);
    if (e != hipSuccess) return fail_hip(e, "distance kernel launch");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_signed_distance(cubvh_bvh* bvh, int device, int64_t n, uint32_t mode, const float* positions,
                          float* distances, int64_t* face_id, float* uvw, void* stream) {
    // src/bvh.cu:440-465: mode 0 is watertight, every other value raystab
    return run_distance(bvh, device, n, mode == 0 ? 1 : 2, positions, distances, face_id, uvw, stream);
}

int cubvh_ray_trace(cubvh_bvh* bvh, int device, int64_t n, const float* rays_o, const float* rays_d,
                    float* positions, int64_t* face_id, float* depth, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!bvh) return fail(CUBVH_ERR_ARG, "bvh is null");
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative ray count");
    if (n == 0) return CUBVH_OK;
    if (!rays_o || !rays_d || !positions || !face_id || !depth) return fail(CUBVH_ERR_ARG, "null ray buffer");
    if (n > ((int64_t)1 << 37)) return fail(CUBVH_ERR_ARG, "too many rays (at most 2^37 per call)");
    int rc = bvh->ensure_device(device, (hipStream_t)stream);
    if (rc != CUBVH_OK) return rc;
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    QueryArgs q = bvh->args(device, n, stream);
    q.in0 = rays_o;
    q.in1 = rays_d;
    q.out_v = positions;
    q.out_id = face_id;
    q.out_f = depth;
    // packets from 2,048 rays on, walked in an order that keeps rays of close origins and
    // directions together (stream-ordered scratch, released after the launch)
    q.packet = n >= kMortonMin;
    hipStream_t hs = (hipStream_t)stream;
    void* scratch = nullptr;
    if (q.packet && n < ((int64_t)1 << 31)) {
        rc = ensure_pool(device);
        if (rc != CUBVH_OK) return rc;
        const size_t ob = ((size_t)n * 4 + 255) & ~(size_t)255;
        hipError_t e = hipMallocAsync(&scratch, ob + morton_workspace_bytes(n), hs);
        if (e != hipSuccess) return fail_hip(e, "hipMallocAsync(ray order)");
        e = ray_order(n, rays_o, rays_d, (char*)scratch + ob, (uint32_t*)scratch, hs);
        if (e != hipSuccess) {
            (void)hipFreeAsync(scratch, hs);
            return fail_hip(e, "ray ordering");
        }
        q.order = (const uint32_t*)scratch;
    }
    hipError_t e = launch_ray(q);
    if (scratch) {
        hipError_t e2 = hipFreeAsync(scratch, hs);
        if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return fail_hip(e, "ray kernel launch");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_floodfill(int device, const uint8_t* grid, int32_t B, int32_t H, int32_t W, int32_t D, int32_t* labels,
                    void* stream) {
    CUBVH_GUARD_BEGIN
    if (B < 0 || H < 0 || W < 0 || D < 0) return fail(CUBVH_ERR_ARG, "negative grid dimension");
    const int64_t ntot = (int64_t)B * H * W * D;
    if (ntot == 0) return CUBVH_OK;
    if (ntot >= ((int64_t)1 << 31)) return fail(CUBVH_ERR_ARG, "grid must have fewer than 2^31 voxels");
    if (!grid || !labels) return fail(CUBVH_ERR_ARG, "null grid buffer");
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    hipError_t e = launch_floodfill(grid, B, H, W, D, labels, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "floodfill kernel launch");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

namespace {

int udf_backward_impl(int device, int64_t n, const float* grad_dist, const float* positions, const int64_t* face_id,
                      const float* uvw, const float* vertices, int64_t n_vertices, const int32_t* faces, int64_t n_faces,
                      float* grad_points, float* grad_vertices, const uint32_t* corners, const uint32_t* corner_begin,
                      const uint32_t* corner_end, void* stream) {
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative query count");
    if (n == 0) return CUBVH_OK;
    if (!grad_dist || !positions || !face_id || !uvw || !vertices || !faces)
        return fail(CUBVH_ERR_ARG, "null input buffer");
    if (n_vertices <= 0 || n_faces <= 0) return fail(CUBVH_ERR_ARG, "empty mesh");
    // the vertex gradients sort queries by face and face corners by vertex (int counts)
    if (n >= ((int64_t)1 << 31)) return fail(CUBVH_ERR_ARG, "too many queries (at most 2^31 - 1 per call)");
    if (n_vertices >= ((int64_t)1 << 31) || 3 * n_faces >= ((int64_t)1 << 31))
        return fail(CUBVH_ERR_ARG, "mesh too large for the backward (vertices and 3 x faces below 2^31)");
    if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    hipStream_t hs = (hipStream_t)stream;
    void* ws = nullptr;
    if (grad_vertices) {
        int rc = ensure_pool(device);
        if (rc != CUBVH_OK) return rc;
        hipError_t e = hipMallocAsync(
            &ws, udf_backward_workspace_bytes(n, n_vertices, n_faces, grad_points != nullptr, corners != nullptr), hs);
        if (e != hipSuccess) return fail_hip(e, "hipMallocAsync(backward workspace)");
    }
    hipError_t e = launch_udf_backward(n, grad_dist, positions, face_id, uvw, vertices, n_vertices, faces, n_faces,
                                       grad_points, grad_vertices, corners, corner_begin, corner_end, ws, hs);
    if (ws) {
        hipError_t e2 = hipFreeAsync(ws, hs);
        if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return fail_hip(e, "backward kernel launch");
    return CUBVH_OK;
}

}  // namespace

int cubvh_udf_backward(int device, int64_t n, const float* grad_dist, const float* positions,
                       const int64_t* face_id, const float* uvw, const float* vertices, int64_t n_vertices,
                       const int32_t* faces, int64_t n_faces, float* grad_points, float* grad_vertices,
                       void* stream) {
    CUBVH_GUARD_BEGIN
    return udf_backward_impl(device, n, grad_dist, positions, face_id, uvw, vertices, n_vertices, faces, n_faces,
                             grad_points, grad_vertices, nullptr, nullptr, nullptr, stream);
    CUBVH_GUARD_END
}

int cubvh_udf_backward_with_corners(int device, int64_t n, const float* grad_dist, const float* positions,
                                    const int64_t* face_id, const float* uvw, const float* vertices,
                                    int64_t n_vertices, const int32_t* faces, int64_t n_faces, float* grad_points,
                                    float* grad_vertices, const uint32_t* corners, const uint32_t* corner_begin,
                                    const uint32_t* corner_end, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!corners || !corner_begin || !corner_end) return fail(CUBVH_ERR_ARG, "null vertex corner table");
    return udf_backward_impl(device, n, grad_dist, positions, face_id, uvw, vertices, n_vertices, faces, n_faces,
                             grad_points, grad_vertices, corners, corner_begin, corner_end, stream);
    CUBVH_GUARD_END
}

int cubvh_vertex_corners(int device, const int32_t* faces, int64_t n_faces, int64_t n_vertices, uint32_t* corners,
                         uint32_t* corner_begin, uint32_t* corner_end, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!faces || !corners || !corner_begin || !corner_end) return fail(CUBVH_ERR_ARG, "null buffer");
    if (n_vertices <= 0 || n_faces <= 0) return fail(CUBVH_ERR_ARG, "empty mesh");
    if (n_vertices >= ((int64_t)1 << 31) || 3 * n_faces >= ((int64_t)1 << 31))
        return fail(CUBVH_ERR_ARG, "mesh too large (vertices and 3 x faces below 2^31)");
    if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    int rc = ensure_pool(device);
    if (rc != CUBVH_OK) return rc;
    hipStream_t hs = (hipStream_t)stream;
    void* ws = nullptr;
    hipError_t e = hipMallocAsync(&ws, vertex_corners_workspace_bytes(n_vertices, n_faces), hs);
    if (e != hipSuccess) return fail_hip(e, "hipMallocAsync(corner sort workspace)");
    e = launch_vertex_corners(faces, n_faces, n_vertices, corners, corner_begin, corner_end, ws, hs);
    hipError_t e2 = hipFreeAsync(ws, hs);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail_hip(e, "vertex corner table");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

namespace {

// Forest of trees that all live on `device` only (GPU-built): their device arrays are
// copied into one block there, node references rebased by a kernel.
int forest_create_on_device(cubvh_bvh* const* bvhs, int32_t n_meshes, int32_t device, cubvh_forest** out) {
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    cubvh_forest* f = new cubvh_forest;
    f->n_meshes = n_meshes;
    f->resident_device = device;
    f->mesh_info.assign(4 * (size_t)n_meshes, 0);
    for (int32_t m = 0; m < n_meshes; ++m) {
        const cubvh_bvh* b = bvhs[m];
        if (f->n_faces + b->host.n_faces > MAX_LEAF_FIRST || f->n_nodes + b->n_nodes > 0x7fffffff) {
            delete f;
            return fail(CUBVH_ERR_ARG, "forest too large for the 27-bit leaf encoding");
        }
        f->mesh_info[4 * m + 0] = (int32_t)f->n_nodes;
        f->mesh_info[4 * m + 1] = (int32_t)(f->n_faces + b->host.face0_slot);
        f->mesh_info[4 * m + 2] = (int32_t)b->host.n_faces;
        f->n_nodes += b->n_nodes;
        f->n_faces += b->host.n_faces;
        f->stack_need = std::max(f->stack_need, b->host.stack_need);
    }
    if (stack_capacity_for(f->stack_need) < 0) {
        delete f;
        return fail(CUBVH_ERR_STATE, "tree too deep for the traversal stack");
    }
    cubvh_forest::DevCopy& c = f->dev[device];
    hipError_t e = hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = c.tree.reserve(f->n_nodes, f->n_faces);
    const size_t mb = f->mesh_info.size() * sizeof(int32_t);
    if (e == hipSuccess) e = hipMalloc(&c.mesh_info, mb);
    if (e == hipSuccess) e = hipMemcpy(c.mesh_info, f->mesh_info.data(), mb, hipMemcpyHostToDevice);
    const hipStream_t s = nullptr;
    int64_t node_base = 0, tri_base = 0;
    for (int32_t m = 0; m < n_meshes && e == hipSuccess; ++m) {
        const cubvh_bvh* b = bvhs[m];
        const TreeArrays& src = b->dev[device].view;
        const TreeArrays& dst = c.tree.view;
        const int64_t nf = b->host.n_faces, nn = b->n_nodes;
        auto cp = [&](const void* to, const void* from, size_t bytes) {
            if (e == hipSuccess && bytes) e = hipMemcpyAsync(const_cast<void*>(to), from, bytes, hipMemcpyDeviceToDevice, s);
        };
        cp(dst.nodes + node_base, src.nodes, (size_t)nn * sizeof(WideNode));
        cp(dst.tris + 3 * tri_base, src.tris, (size_t)nf * 3 * sizeof(float4));
        cp(dst.tbnd + 2 * tri_base, src.tbnd, (size_t)nf * 2 * sizeof(float4));
        cp(dst.lbnd + 2 * tri_base, src.lbnd, (size_t)nf * 2 * sizeof(float4));
        cp(dst.cbnd + 8 * node_base, src.cbnd, (size_t)nn * 8 * sizeof(float4));
        cp(dst.tpre + 9 * tri_base, src.tpre, (size_t)nf * 9 * sizeof(float4));
        if (e == hipSuccess)
            e = launch_rebase_nodes(const_cast<WideNode*>(dst.nodes) + node_base, nn, (int32_t)node_base,
                                    (int32_t)tri_base, s);
        node_base += nn;
        tri_base += nf;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        delete f;
        return fail_hip(e, "forest assembly on the device");
    }
    *out = f;
    return CUBVH_OK;
}

}  // namespace

int cubvh_forest_create(cubvh_bvh* const* bvhs, int32_t n_meshes, cubvh_forest** out) {
    CUBVH_GUARD_BEGIN
    if (!out) return fail(CUBVH_ERR_ARG, "out is null");
    *out = nullptr;
    if (!bvhs || n_meshes <= 0) return fail(CUBVH_ERR_ARG, "need at least one BVH");
    for (int32_t m = 0; m < n_meshes; ++m)
        if (!bvhs[m]) return fail(CUBVH_ERR_ARG, "null BVH in the list");
    // every tree built on one device and not copied to the host: assemble there
    const int32_t d0 = bvhs[0]->mirror_device;
    bool on_device = d0 >= 0;
    for (int32_t m = 1; m < n_meshes && on_device; ++m) on_device = bvhs[m]->mirror_device == d0;
    if (on_device) return forest_create_on_device(bvhs, n_meshes, d0, out);
    cubvh_forest* f = new cubvh_forest;
    f->n_meshes = n_meshes;
    f->mesh_info.assign(4 * (size_t)n_meshes, 0);
    for (int32_t m = 0; m < n_meshes; ++m) {
        int rc = bvhs[m]->ensure_host_derived();
        if (rc != CUBVH_OK) {
            delete f;
            return rc;
        }
        const HostBvh& h = bvhs[m]->host;
        const int64_t node_base = (int64_t)f->nodes.size();
        const int64_t tri_base = f->n_faces;
        if (tri_base + h.n_faces > MAX_LEAF_FIRST || node_base + (int64_t)h.nodes.size() > 0x7fffffff) {
            delete f;
            return fail(CUBVH_ERR_ARG, "forest too large for the 27-bit leaf encoding");
        }
        for (WideNode w : h.nodes) {
            for (int k = 0; k < 4; ++k) {
                int32_t r = w.child[k];
                if (r == CHILD_EMPTY) continue;
                if (r >= 0) {
                    w.child[k] = (int32_t)(r + node_base);
                } else {
                    w.child[k] = make_leaf_ref((int64_t)leaf_ref_first(r) + tri_base, leaf_ref_count(r));
                }
            }
            f->nodes.push_back(w);
        }
        f->tris.insert(f->tris.end(), h.tris.begin(), h.tris.end());
        f->tbnd.insert(f->tbnd.end(), h.tbnd.begin(), h.tbnd.end());
        f->lbnd.insert(f->lbnd.end(), h.lbnd.begin(), h.lbnd.end());
        f->cbnd.insert(f->cbnd.end(), h.cbnd.begin(), h.cbnd.end());
        f->tpre.insert(f->tpre.end(), h.tpre.begin(), h.tpre.end());
        f->mesh_info[4 * m + 0] = (int32_t)node_base;
        f->mesh_info[4 * m + 1] = (int32_t)(tri_base + h.face0_slot);
        f->mesh_info[4 * m + 2] = (int32_t)h.n_faces;
        f->n_faces += h.n_faces;
        f->stack_need = std::max(f->stack_need, h.stack_need);
    }
    f->n_nodes = (int64_t)f->nodes.size();
    if (stack_capacity_for(f->stack_need) < 0) {
        delete f;
        return fail(CUBVH_ERR_STATE, "tree too deep for the traversal stack");
    }
    *out = f;
    return CUBVH_OK;
    CUBVH_GUARD_END
}

void cubvh_forest_destroy(cubvh_forest* forest) { delete forest; }

int cubvh_forest_info(const cubvh_forest* forest, int64_t info[4]) {
    if (!forest || !info) return fail(CUBVH_ERR_ARG, "null argument");
    info[0] = forest->n_meshes;
    info[1] = forest->n_nodes;
    info[2] = forest->n_faces;
    info[3] = forest->stack_need;
    return CUBVH_OK;
}

// Largest Morton-bin count of a forest call (n_meshes << order bits): bin indices are
// 32-bit and the bin histogram takes 4 bytes per bin.
constexpr int64_t kMaxForestBins = (int64_t)1 << 30;

int cubvh_forest_reserve(cubvh_forest* forest, int device, int64_t max_queries, int32_t max_order_bits,
                         void* stream) {
    CUBVH_GUARD_BEGIN
    if (!forest) return fail(CUBVH_ERR_ARG, "forest is null");
    if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    if (max_queries < 0 || max_queries >= ((int64_t)1 << 32) - 64 * (int64_t)forest->n_meshes)
        return fail(CUBVH_ERR_ARG, "max_queries out of range");
    if (max_order_bits < 0 || max_order_bits > 16) return fail(CUBVH_ERR_ARG, "order bits must be in 0..16");
    if (((int64_t)forest->n_meshes << max_order_bits) > kMaxForestBins)
        return fail(CUBVH_ERR_ARG, "n_meshes * 2^order_bits exceeds 2^30 Morton bins: use fewer order bits");
    std::lock_guard<std::mutex> lock(forest->mu);
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    int rc = forest->upload(device, (hipStream_t)stream);
    if (rc != CUBVH_OK) return rc;
    return forest->reserve(device, max_queries, max_order_bits, (hipStream_t)stream);
    CUBVH_GUARD_END
}

int cubvh_forest_unsigned_distance(cubvh_forest* forest, int device, int64_t n, const float* positions,
                                   const int64_t* query_offsets, float* distances, int64_t* face_id, float* uvw,
                                   const int32_t* tuning, int32_t n_tuning, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!forest) return fail(CUBVH_ERR_ARG, "forest is null");
    if (n_tuning < 0 || n_tuning > 5 || (n_tuning > 0 && !tuning))
        return fail(CUBVH_ERR_ARG, "tuning must hold 0 to 5 entries");
    if (device < 0 || device >= kMaxDevices) return fail(CUBVH_ERR_ARG, "device ordinal out of range");
    if (n < 0 || n >= ((int64_t)1 << 32) - 64 * (int64_t)forest->n_meshes)
        return fail(CUBVH_ERR_ARG, "query count out of range");
    if (n == 0) return CUBVH_OK;
    if (!positions || !query_offsets || !distances || !face_id) return fail(CUBVH_ERR_ARG, "null query buffer");
    const int32_t bits = n_tuning > 0 ? tuning[0] : 12;
    const int32_t persistent = n_tuning > 1 ? tuning[1] : 1;
    const int32_t wpc = n_tuning > 2 ? tuning[2] : 8;
    const int32_t wpb = n_tuning > 3 ? tuning[3] : 4;
    const int32_t xcd = n_tuning > 4 ? tuning[4] : 0;
    if (wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8) return fail(CUBVH_ERR_ARG, "waves per block must be 1, 2, 4 or 8");
    if (xcd < 0 || xcd > 65536) return fail(CUBVH_ERR_ARG, "XCD chunk must be in 0..65536 workgroups");
    if (bits < 0 || bits > 16) return fail(CUBVH_ERR_ARG, "order bits must be in 0..16");
    // a bin index is mesh << bits | code in 32 bits; the histogram holds one word per bin
    if (((int64_t)forest->n_meshes << bits) > kMaxForestBins)
        return fail(CUBVH_ERR_ARG, "n_meshes * 2^order_bits exceeds 2^30 Morton bins: use fewer order bits");
    if (wpc < 1 || wpc > 32) return fail(CUBVH_ERR_ARG, "waves per CU must be in 1..32");
    hipStream_t hs = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(forest->mu);
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    cubvh_forest::DevCopy& c = forest->dev[device];
    if (!c.tree.block || !c.ws || c.ws_n < n || c.ws_bits < bits) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hipError_t e = hipStreamIsCapturing(hs, &cs);
        if (e != hipSuccess) return fail_hip(e, "hipStreamIsCapturing");
        if (cs != hipStreamCaptureStatusNone)
            return fail(CUBVH_ERR_ARG, "forest workspace too small while capturing: call cubvh_forest_reserve first");
        int rc = forest->upload(device, hs);
        if (rc == CUBVH_OK) rc = forest->reserve(device, n, bits, hs);
        if (rc != CUBVH_OK) return rc;
    }
    ForestWorkspace ws = carve_forest_workspace(c.ws, c.ws_n, forest->n_meshes, c.ws_bits);
    // bins actually used by this call
    ws.nbins = (int64_t)forest->n_meshes << bits;
    ws.nchunks = (ws.nbins + 1023) / 1024;
    hipError_t e = forest_order(n, positions, query_offsets, forest->n_meshes, bits, ws, hs);
    if (e != hipSuccess) return fail_hip(e, "forest query ordering");
    ForestArgs a;
    a.n = n;
    a.n_meshes = forest->n_meshes;
    a.positions = positions;
    a.offsets = query_offsets;
    a.distances = distances;
    a.face_id = face_id;
    a.uvw = uvw;
    a.tree = c.tree.view;
    a.mesh_info = c.mesh_info;
    a.persistent = persistent;
    a.waves_per_cu = wpc;
    a.waves_per_block = wpb;
    a.xcd_chunk = xcd;
    a.n_cu = c.n_cu;
    e = launch_forest_distance(a, ws, hs);
    if (e != hipSuccess) return fail_hip(e, "forest kernel launch");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

// ---------------------------------------------------------------------------
// Sparse-voxel hash table (include/gpu/hashtable.cuh, src/api_gpu.cu:215-281).

}  // extern "C"

struct cubvh_hashtable {
    std::mutex mu;
    int32_t dims = 3;        // key words expected by the next insert / build
    int32_t table_dims = 3;  // key words the table memory is laid out for
    int64_t capacity = 0;
    int device = -1;
    int32_t* table = nullptr;
    size_t bytes = 0;
    int64_t rows = 0;  // rows inserted since the table was last cleared

    // Memory for `cap` slots of keys of `d` words on `dev`, every slot empty.
    int lay_out(int dev, int64_t cap, int32_t d, hipStream_t s) {
        const size_t need = (size_t)cap * (size_t)hash_slot_words(d) * sizeof(int32_t);
        DeviceScope g(dev);
        if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
        if (table && (device != dev || bytes < need)) {
            int rc = release();
            if (rc != CUBVH_OK) return rc;
        }
        if (!table && need > 0) {
            hipError_t e = hipMalloc((void**)&table, need);
            if (e != hipSuccess) {
                table = nullptr;
                return fail_hip(e, "hipMalloc (hash table)");
            }
            bytes = need;
        }
        device = dev;
        capacity = cap;
        table_dims = d;
        rows = 0;
        hipError_t e = hash_clear(table, cap, hash_slot_words(d), s);
        if (e != hipSuccess) return fail_hip(e, "hash table clear");
        return CUBVH_OK;
    }
    int release() {
        if (!table) return CUBVH_OK;
        DeviceScope g(device);
        hipError_t e = g.err == hipSuccess ? hipFree(table) : g.err;
        table = nullptr;
        bytes = 0;
        if (e != hipSuccess) return fail_hip(e, "hipFree (hash table)");
        return CUBVH_OK;
    }
    int check_device(int dev) const {
        if (table && dev != device)
            return fail(CUBVH_ERR_ARG, "the hash table lives on device " + std::to_string(device) +
                                           ", the keys on device " + std::to_string(dev));
        return CUBVH_OK;
    }
    int insert(int dev, const int32_t* coords, int64_t n, int32_t d, hipStream_t s) {
        if (capacity <= 0 || n <= 0) {  // as the reference: nothing to insert
            dims = d;
            return CUBVH_OK;
        }
        if (!coords) return fail(CUBVH_ERR_ARG, "null coords buffer");
        if (n > (int64_t)INT32_MAX) return fail(CUBVH_ERR_ARG, "at most 2^31 - 1 rows per insert");
        int rc = check_device(dev);
        if (rc != CUBVH_OK) return rc;
        if (d != table_dims && rows > 0)
            return fail(CUBVH_ERR_ARG, "the table holds keys of " + std::to_string(table_dims) +
                                           " words; prepare() it before inserting keys of " + std::to_string(d));
        dims = d;
        if (d != table_dims) {
            rc = lay_out(dev, capacity, d, s);
            if (rc != CUBVH_OK) return rc;
        }
        DeviceScope g(dev);
        if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
        hipError_t e = hash_insert(table, capacity, hash_slot_words(d), coords, n, d, s);
        if (e != hipSuccess) return fail_hip(e, "hash insert kernel launch");
        rows += n;
        return CUBVH_OK;
    }
};

namespace {
int check_dims(int32_t d) {
    if (d < 1 || d > kHashMaxDims)
        return fail(CUBVH_ERR_ARG, "keys must have 1.." + std::to_string(kHashMaxDims) + " words, got " +
                                       std::to_string(d));
    return CUBVH_OK;
}
}  // namespace

extern "C" {

int cubvh_hashtable_create(int32_t num_dims, cubvh_hashtable** out) {
    CUBVH_GUARD_BEGIN
    if (!out) return fail(CUBVH_ERR_ARG, "null output pointer");
    *out = nullptr;
    int rc = check_dims(num_dims);
    if (rc != CUBVH_OK) return rc;
    cubvh_hashtable* h = new cubvh_hashtable();
    h->dims = h->table_dims = num_dims;
    *out = h;
    return CUBVH_OK;
    CUBVH_GUARD_END
}

void cubvh_hashtable_destroy(cubvh_hashtable* ht) {
    if (!ht) return;
    (void)ht->release();
    delete ht;
}

int cubvh_hashtable_set_num_dims(cubvh_hashtable* ht, int32_t num_dims) {
    CUBVH_GUARD_BEGIN
    if (!ht) return fail(CUBVH_ERR_ARG, "null hash table");
    int rc = check_dims(num_dims);
    if (rc != CUBVH_OK) return rc;
    std::lock_guard<std::mutex> lock(ht->mu);
    ht->dims = num_dims;
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_hashtable_info(cubvh_hashtable* ht, int64_t info[5]) {
    CUBVH_GUARD_BEGIN
    if (!ht || !info) return fail(CUBVH_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lock(ht->mu);
    info[0] = ht->dims;
    info[1] = ht->capacity;
    info[2] = ht->table_dims;
    info[3] = hash_slot_words(ht->table_dims);
    info[4] = ht->table ? ht->device : -1;
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_hashtable_resize(cubvh_hashtable* ht, int device, int64_t capacity, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!ht) return fail(CUBVH_ERR_ARG, "null hash table");
    if (capacity < 0 || capacity > (int64_t)INT32_MAX) return fail(CUBVH_ERR_ARG, "capacity must be in [0, 2^31 - 1]");
    std::lock_guard<std::mutex> lock(ht->mu);
    if (capacity == 0) {
        ht->capacity = 0;
        ht->rows = 0;
        return ht->release();
    }
    return ht->lay_out(device, capacity, ht->dims, (hipStream_t)stream);
    CUBVH_GUARD_END
}

int cubvh_hashtable_prepare(cubvh_hashtable* ht, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!ht) return fail(CUBVH_ERR_ARG, "null hash table");
    std::lock_guard<std::mutex> lock(ht->mu);
    if (ht->capacity <= 0) return CUBVH_OK;  // as the reference: nothing to clear
    return ht->lay_out(ht->device, ht->capacity, ht->dims, (hipStream_t)stream);
    CUBVH_GUARD_END
}

int cubvh_hashtable_insert(cubvh_hashtable* ht, int device, const int32_t* coords, int64_t n, int32_t num_dims,
                           void* stream) {
    CUBVH_GUARD_BEGIN
    if (!ht) return fail(CUBVH_ERR_ARG, "null hash table");
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative row count");
    int rc = check_dims(num_dims);
    if (rc != CUBVH_OK) return rc;
    std::lock_guard<std::mutex> lock(ht->mu);
    return ht->insert(device, coords, n, num_dims, (hipStream_t)stream);
    CUBVH_GUARD_END
}

int cubvh_hashtable_build(cubvh_hashtable* ht, int device, const int32_t* coords, int64_t n, int32_t num_dims,
                          void* stream) {
    CUBVH_GUARD_BEGIN
    if (!ht) return fail(CUBVH_ERR_ARG, "null hash table");
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative row count");
    if (n > ((int64_t)INT32_MAX - 1) / 2) return fail(CUBVH_ERR_ARG, "at most 2^30 - 1 rows per build");
    if (n > 0 && !coords) return fail(CUBVH_ERR_ARG, "null coords buffer");
    int rc = check_dims(num_dims);
    if (rc != CUBVH_OK) return rc;
    std::lock_guard<std::mutex> lock(ht->mu);
    ht->dims = num_dims;
    rc = ht->lay_out(device, std::max<int64_t>(16, 2 * n), num_dims, (hipStream_t)stream);
    if (rc != CUBVH_OK) return rc;
    return ht->insert(device, coords, n, num_dims, (hipStream_t)stream);
    CUBVH_GUARD_END
}

int cubvh_hashtable_search(cubvh_hashtable* ht, int device, const int32_t* queries, int64_t m, int32_t num_dims,
                           int32_t* out, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!ht) return fail(CUBVH_ERR_ARG, "null hash table");
    if (m < 0) return fail(CUBVH_ERR_ARG, "negative query count");
    std::lock_guard<std::mutex> lock(ht->mu);
    if (ht->capacity <= 0 || !ht->table) return fail(CUBVH_ERR_ARG, "hash table is not built");
    if (num_dims != ht->table_dims)
        return fail(CUBVH_ERR_ARG, "queries have " + std::to_string(num_dims) + " words, the table's keys " +
                                       std::to_string(ht->table_dims));
    if (m == 0) return CUBVH_OK;
    if (!queries || !out) return fail(CUBVH_ERR_ARG, "null query or output buffer");
    int rc = ht->check_device(device);
    if (rc != CUBVH_OK) return rc;
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    hipError_t e = hash_search(ht->table, ht->capacity, hash_slot_words(num_dims), queries, m, num_dims, out,
                               (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "hash search kernel launch");
    return CUBVH_OK;
    CUBVH_GUARD_END
}

// ---------------------------------------------------------------------------
// Sparse marching cubes (include/gpu/spcumc.cuh:631-868, src/api_gpu.cu:154-211).

}  // extern "C"

struct cubvh_smc {
    SmcState* st = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n_vertices = 0, n_triangles = 0;
    bool written = false;
};

extern "C" {

int cubvh_sparse_marching_cubes(int device, const int32_t* coords, const float* corners, int64_t n, float iso,
                                int32_t ensure_consistency, cubvh_smc** out, int64_t* n_vertices,
                                int64_t* n_triangles, void* stream) {
    CUBVH_GUARD_BEGIN
    if (!out || !n_vertices || !n_triangles) return fail(CUBVH_ERR_ARG, "null output pointer");
    *out = nullptr;
    *n_vertices = *n_triangles = 0;
    if (n < 0) return fail(CUBVH_ERR_ARG, "negative cell count");
    if (n > kSmcMaxCells) return fail(CUBVH_ERR_ARG, "at most 2^28 cells per extraction");
    if (n > 0 && (!coords || !corners)) return fail(CUBVH_ERR_ARG, "null coords or corners buffer");
    DeviceScope g(device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    int rc = ensure_pool(device);
    if (rc != CUBVH_OK) return rc;
    SmcState* st = nullptr;
    int64_t nv = 0, nt = 0;
    SmcStatus r = smc_extract(coords, corners, n, iso, ensure_consistency != 0, (hipStream_t)stream, &st, &nv, &nt);
    if (r.code != 0) return fail(r.code == -1 ? CUBVH_ERR_ARG : CUBVH_ERR_HIP, r.message);
    cubvh_smc* h = new cubvh_smc();
    h->st = st;
    h->device = device;
    h->stream = (hipStream_t)stream;
    h->n_vertices = nv;
    h->n_triangles = nt;
    *out = h;
    *n_vertices = nv;
    *n_triangles = nt;
    return CUBVH_OK;
    CUBVH_GUARD_END
}

int cubvh_smc_write(cubvh_smc* smc, float* vertices, int32_t* triangles) {
    CUBVH_GUARD_BEGIN
    if (!smc || !smc->st) return fail(CUBVH_ERR_ARG, "null or released extraction");
    if (smc->written) return fail(CUBVH_ERR_ARG, "the extraction was already written");
    if ((smc->n_vertices > 0 && !vertices) || (smc->n_triangles > 0 && !triangles))
        return fail(CUBVH_ERR_ARG, "null output buffer");
    DeviceScope g(smc->device);
    if (g.err != hipSuccess) return fail_hip(g.err, "hipSetDevice");
    SmcStatus r = smc_write(smc->st, vertices, triangles, smc->stream);
    if (r.code != 0) return fail(CUBVH_ERR_HIP, r.message);
    smc->written = true;
    return CUBVH_OK;
    CUBVH_GUARD_END
}

void cubvh_smc_destroy(cubvh_smc* smc) {
    if (!smc) return;
    if (smc->st) {
        DeviceScope g(smc->device);
        smc_release(smc->st, smc->stream);
    }
    delete smc;
}

}  // extern "C"
