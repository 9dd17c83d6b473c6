// bindings.cpp -- the `_cubvh` extension module: a drop-in for the reference package's
// native module (src/bindings.cpp:14-87, over the interfaces of include/gpu/api_gpu.h:17-53),
// built on this library's C ABI (include/cubvh_amd.h).  The reference's Python layer
// (cubvh/api.py) imports it as `import _cubvh as _backend` and calls:
//
//   _backend.create_cuBVH(vertices, triangles)                     cubvh/api.py:24
//   _backend.from_state_dict(state_dict)                           cubvh/api.py:33
//   impl.state_dict()                                              cubvh/api.py:28
//   impl.ray_trace(rays_o, rays_d, positions, face_id, depth)      cubvh/api.py:59
//   impl.unsigned_distance(positions, distances, face_id, uvw)     cubvh/api.py:88
//   impl.signed_distance(positions, distances, face_id, uvw, mode) cubvh/api.py:119
//   _backend.floodfill(grid)                                       cubvh/api.py:136-138
//   _backend.create_cuHashTable() and its methods                  cubvh/api.py:153-169
//   _backend.sparse_marching_cubes(coords, corners, iso, ensure)   cubvh/api.py:185
//
// Semantics kept from the reference binding:
//  * create_cuBVH takes host arrays and force-casts them as the reference's pybind11
//    Eigen caster does (patch/eigen/matrix.h:490-524): vertices to float32, triangles to
//    uint32 (numpy casting rules), both C-contiguous [N, 3]; any other shape is a
//    TypeError.  Construction is host-only (no GPU touched until the first query).
//  * query methods write into caller-allocated tensors on torch's current stream and
//    return without synchronizing; uvw may be None; mode 0 = watertight, other = raystab.
//  * state_dict() returns the reference's own layout: {"nodes": int32 [n, 9],
//    "triangles": int32 [M, 12]} (bvh.cuh:84-94, api_gpu.cu:110-121); from_state_dict
//    accepts that layout (and this library's [n, 32] node layout), with the reference's
//    errors for a missing key and for triangles not on the CPU (api_gpu.cu:47-56).
//  * floodfill takes a bool [B, H, W, D] grid and returns int32 labels of the same shape.
//  * the CPU-only names of the reference module (HashTable, fill_holes, merge_vertices,
//    sparse_marching_cubes_cpu, decimate, parallel_decimate) exist and raise
//    NotImplementedError: they are outside the BVH query path.
// Stricter than the reference: query tensors are checked (GPU, device, dtype,
// contiguity, size) before any launch, where the reference reads them unchecked.
#include <torch/extension.h>

#include <c10/hip/HIPFunctions.h>
#include <c10/hip/HIPStream.h>
#include <pybind11/numpy.h>

#include <string>

#include "cubvh_amd.h"

namespace py = pybind11;

namespace {

void check_rc(int rc, const char* what) {
    if (rc != CUBVH_OK) throw std::runtime_error(std::string(what) + ": " + cubvh_last_error());
}

void* current_stream(const at::Tensor& t) {
    return (void*)c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

// A query-side tensor: on the GPU of `like`, of type dt, contiguous, holding at least
// n * cols elements.
void check_tensor(const at::Tensor& t, const char* name, at::ScalarType dt, int64_t n, int64_t cols,
                  const at::Tensor& like) {
    TORCH_CHECK(t.is_cuda(), name, " must reside on the GPU");
    TORCH_CHECK(t.device() == like.device(), name, " must be on the device of the queries");
    TORCH_CHECK(t.scalar_type() == dt, name, " must be ", c10::toString(dt), ", got ", t.scalar_type());
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
    TORCH_CHECK(t.numel() >= n * cols, name, " holds ", t.numel(), " elements, ", n * cols, " needed");
}

// n = size(0) of a [N, 3] query tensor, as the reference reads it (api_gpu.cu:72,85,99).
int64_t rows_of(const at::Tensor& queries, const char* name) {
    TORCH_CHECK(queries.dim() == 2 && queries.size(1) == 3, name, " must be [N, 3], got ", queries.sizes());
    return queries.size(0);
}

[[noreturn]] void not_implemented(const std::string& name) {
    PyErr_SetString(PyExc_NotImplementedError,
                    (name + " is not implemented by this module (outside the BVH query path)").c_str());
    throw py::error_already_set();
}

// The impl object returned by create_cuBVH / from_state_dict (api_gpu.h:17-26).
class BvhImpl {
public:
    explicit BvhImpl(cubvh_bvh* h) : h_(h) {}
    ~BvhImpl() {
        if (h_) cubvh_bvh_destroy(h_);
    }
    BvhImpl(const BvhImpl&) = delete;
    BvhImpl& operator=(const BvhImpl&) = delete;

    // api_gpu.cu:70-81
    void ray_trace(at::Tensor rays_o, at::Tensor rays_d, at::Tensor positions, at::Tensor face_id, at::Tensor depth) {
        const int64_t n = rows_of(rays_o, "rays_o");
        check_tensor(rays_o, "rays_o", at::kFloat, n, 3, rays_o);
        check_tensor(rays_d, "rays_d", at::kFloat, n, 3, rays_o);
        check_tensor(positions, "positions", at::kFloat, n, 3, rays_o);
        check_tensor(face_id, "face_id", at::kLong, n, 1, rays_o);
        check_tensor(depth, "depth", at::kFloat, n, 1, rays_o);
        check_rc(cubvh_ray_trace(h_, rays_o.device().index(), n, rays_o.data_ptr<float>(), rays_d.data_ptr<float>(),
                                 positions.data_ptr<float>(), face_id.data_ptr<int64_t>(), depth.data_ptr<float>(),
                                 current_stream(rays_o)),
                 "ray_trace");
    }

    // api_gpu.cu:83-95
    void unsigned_distance(at::Tensor positions, at::Tensor distances, at::Tensor face_id,
                           c10::optional<at::Tensor> uvw) {
        const int64_t n = rows_of(positions, "positions");
        float* u = check_outputs(positions, distances, face_id, uvw, n);
        check_rc(cubvh_unsigned_distance(h_, positions.device().index(), n, positions.data_ptr<float>(),
                                         distances.data_ptr<float>(), face_id.data_ptr<int64_t>(), u,
                                         current_stream(positions)),
                 "unsigned_distance");
    }

    // api_gpu.cu:97-108
    void signed_distance(at::Tensor positions, at::Tensor distances, at::Tensor face_id, c10::optional<at::Tensor> uvw,
                         uint32_t mode) {
        const int64_t n = rows_of(positions, "positions");
        float* u = check_outputs(positions, distances, face_id, uvw, n);
        check_rc(cubvh_signed_distance(h_, positions.device().index(), n, mode, positions.data_ptr<float>(),
                                       distances.data_ptr<float>(), face_id.data_ptr<int64_t>(), u,
                                       current_stream(positions)),
                 "signed_distance");
    }

    // api_gpu.cu:110-121, in the reference's own layout
    py::dict state_dict() const {
        int64_t n9 = 0, info[5];
        check_rc(cubvh_bvh_info(h_, info), "state_dict");
        check_rc(cubvh_bvh_export_reference(h_, nullptr, &n9, nullptr), "state_dict");
        at::Tensor nodes = at::empty({n9, 9}, at::dtype(at::kInt));
        at::Tensor tris = at::empty({info[1], 12}, at::dtype(at::kInt));
        check_rc(cubvh_bvh_export_reference(h_, nodes.data_ptr<int32_t>(), &n9, tris.data_ptr<int32_t>()),
                 "state_dict");
        py::dict d;
        d["nodes"] = nodes;
        d["triangles"] = tris;
        return d;
    }

private:
    float* check_outputs(const at::Tensor& positions, const at::Tensor& distances, const at::Tensor& face_id,
                         const c10::optional<at::Tensor>& uvw, int64_t n) {
        check_tensor(positions, "positions", at::kFloat, n, 3, positions);
        check_tensor(distances, "distances", at::kFloat, n, 1, positions);
        check_tensor(face_id, "face_id", at::kLong, n, 1, positions);
        if (!uvw.has_value()) return nullptr;
        check_tensor(*uvw, "uvw", at::kFloat, n, 3, positions);
        return uvw->data_ptr<float>();
    }

    cubvh_bvh* h_;
};

using VertArray = py::array_t<float, py::array::c_style | py::array::forcecast>;
using TrigArray = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>;

// api_gpu.cu:128-130 (construction, api_gpu.cu:21-44)
BvhImpl* create_bvh(VertArray vertices, TrigArray triangles) {
    if (vertices.ndim() != 2 || vertices.shape(1) != 3)
        throw py::type_error("create_cuBVH: vertices must be a [N, 3] array");
    if (triangles.ndim() != 2 || triangles.shape(1) != 3)
        throw py::type_error("create_cuBVH: triangles must be a [M, 3] array");
    const int64_t nv = vertices.shape(0), nf = triangles.shape(0);
    const uint32_t* t = triangles.data();
    for (int64_t i = 0; i < 3 * nf; ++i)
        if (t[i] >= (uint64_t)nv)
            throw std::runtime_error("create_cuBVH: triangle " + std::to_string(i / 3) + " references vertex " +
                                     std::to_string(t[i]) + " of " + std::to_string(nv));
    // indices below nv < 2^31 read as int32 unchanged
    TORCH_CHECK(nv < ((int64_t)1 << 31), "create_cuBVH: at most 2^31 - 1 vertices");
    cubvh_bvh* h = nullptr;
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = cubvh_bvh_create(vertices.data(), nv, reinterpret_cast<const int32_t*>(t), nf, 0, &h);
    }
    check_rc(rc, "create_cuBVH");
    return new BvhImpl(h);
}

// api_gpu.cu:132-134 (construction from a state dict, api_gpu.cu:47-68)
BvhImpl* from_state_dict(py::dict state_dict) {
    if (!state_dict.contains("triangles")) throw std::runtime_error("state_dict missing 'triangles' tensor");
    if (!state_dict.contains("nodes")) throw std::runtime_error("state_dict missing 'nodes' tensor");
    at::Tensor tris = state_dict["triangles"].cast<at::Tensor>();
    at::Tensor nodes = state_dict["nodes"].cast<at::Tensor>();
    TORCH_CHECK(tris.is_cpu(), "triangles must be on CPU");
    TORCH_CHECK(nodes.is_cpu(), "nodes must be on CPU");
    tris = tris.to(at::kInt).contiguous();
    nodes = nodes.to(at::kInt).contiguous();
    TORCH_CHECK(tris.dim() == 2 && tris.size(1) == 12, "'triangles' must be int32 [M, 12]");
    TORCH_CHECK(nodes.dim() == 2 && (nodes.size(1) == 9 || nodes.size(1) == 32),
                "'nodes' must be int32 [n, 9] (reference layout) or [n, 32]");
    cubvh_bvh* h = nullptr;
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = nodes.size(1) == 9 ? cubvh_bvh_import_reference(nodes.data_ptr<int32_t>(), nodes.size(0),
                                                               tris.data_ptr<int32_t>(), tris.size(0), 0, &h)
                                : cubvh_bvh_import(nodes.data_ptr<int32_t>(), nodes.size(0), tris.data_ptr<int32_t>(),
                                                   tris.size(0), &h);
    }
    check_rc(rc, "from_state_dict");
    return new BvhImpl(h);
}

// api_gpu.cu:136-152
at::Tensor floodfill(at::Tensor grid) {
    TORCH_CHECK(grid.scalar_type() == at::kBool, "floodfill: grid must be bool, got ", grid.scalar_type());
    TORCH_CHECK(grid.is_cuda(), "floodfill: grid must reside on the GPU");
    TORCH_CHECK(grid.dim() == 4, "floodfill: grid must be [B, H, W, D]");
    grid = grid.contiguous();
    for (int k = 0; k < 4; ++k) TORCH_CHECK(grid.size(k) < ((int64_t)1 << 31), "floodfill: grid too large");
    at::Tensor labels = at::empty(grid.sizes(), grid.options().dtype(at::kInt));
    check_rc(cubvh_floodfill(grid.device().index(), reinterpret_cast<const uint8_t*>(grid.data_ptr<bool>()),
                             (int32_t)grid.size(0), (int32_t)grid.size(1), (int32_t)grid.size(2),
                             (int32_t)grid.size(3), labels.data_ptr<int32_t>(), current_stream(grid)),
             "floodfill");
    return labels;
}

// api_gpu.cu:154-211
std::tuple<at::Tensor, at::Tensor> sparse_marching_cubes(at::Tensor coords, at::Tensor corners, double iso,
                                                         bool ensure_consistency) {
    TORCH_CHECK(coords.is_cuda(), "coords must reside on CUDA");
    TORCH_CHECK(corners.is_cuda(), "corners must reside on CUDA");
    TORCH_CHECK(coords.scalar_type() == at::kInt, "coords must be int32");
    TORCH_CHECK(corners.scalar_type() == at::kFloat, "corners must be float32");
    TORCH_CHECK(coords.dim() == 2 && coords.size(1) == 3, "coords must be of shape [N,3]");
    TORCH_CHECK(corners.dim() == 2 && corners.size(1) == 8, "corners must be of shape [N,8]");
    TORCH_CHECK(coords.size(0) == corners.size(0), "coords and corners must have the same first-dim (N)");
    TORCH_CHECK(coords.device() == corners.device(), "coords and corners must be on one device");
    coords = coords.contiguous();
    corners = corners.contiguous();
    cubvh_smc* h = nullptr;
    int64_t nv = 0, nt = 0;
    check_rc(cubvh_sparse_marching_cubes(coords.device().index(), coords.data_ptr<int32_t>(), corners.data_ptr<float>(),
                                         coords.size(0), (float)iso, ensure_consistency ? 1 : 0, &h, &nv, &nt,
                                         current_stream(coords)),
             "sparse_marching_cubes");
    at::Tensor verts, tris;
    try {
        verts = at::empty({nv, 3}, coords.options().dtype(at::kFloat));
        tris = at::empty({nt, 3}, coords.options().dtype(at::kInt));
        check_rc(cubvh_smc_write(h, verts.data_ptr<float>(), tris.data_ptr<int32_t>()), "sparse_marching_cubes");
    } catch (...) {
        cubvh_smc_destroy(h);
        throw;
    }
    cubvh_smc_destroy(h);
    return {verts, tris};
}

// The impl object returned by create_cuHashTable (api_gpu.h:37-50, api_gpu.cu:215-276).
class HashTableImpl {
public:
    HashTableImpl() { check_rc(cubvh_hashtable_create(3, &h_), "create_cuHashTable"); }
    ~HashTableImpl() {
        if (h_) cubvh_hashtable_destroy(h_);
    }
    HashTableImpl(const HashTableImpl&) = delete;
    HashTableImpl& operator=(const HashTableImpl&) = delete;

    void set_num_dims(int d) { check_rc(cubvh_hashtable_set_num_dims(h_, d), "set_num_dims"); }
    int get_num_dims() const {
        int64_t info[5];
        check_rc(cubvh_hashtable_info(h_, info), "get_num_dims");
        return (int)info[0];
    }
    void resize(int capacity) {
        const int dev = c10::hip::current_device();
        check_rc(cubvh_hashtable_resize(h_, dev, capacity, (void*)c10::hip::getCurrentHIPStream(dev).stream()),
                 "resize");
    }
    void prepare() {
        int64_t info[5];
        check_rc(cubvh_hashtable_info(h_, info), "prepare");
        const int dev = info[4] >= 0 ? (int)info[4] : c10::hip::current_device();
        check_rc(cubvh_hashtable_prepare(h_, (void*)c10::hip::getCurrentHIPStream(dev).stream()), "prepare");
    }
    void insert(at::Tensor coords) {
        coords = checked_keys(coords, "coords");
        check_rc(cubvh_hashtable_insert(h_, coords.device().index(), coords.data_ptr<int32_t>(), coords.size(0),
                                        (int32_t)coords.size(1), current_stream(coords)),
                 "insert");
    }
    void build(at::Tensor coords) {
        coords = checked_keys(coords, "coords");
        check_rc(cubvh_hashtable_build(h_, coords.device().index(), coords.data_ptr<int32_t>(), coords.size(0),
                                       (int32_t)coords.size(1), current_stream(coords)),
                 "build");
    }
    at::Tensor search(at::Tensor queries) const {
        queries = checked_keys(queries, "queries");
        int64_t info[5];
        check_rc(cubvh_hashtable_info(h_, info), "search");
        TORCH_CHECK(info[1] > 0, "hash table is not built");
        at::Tensor out = at::empty({queries.size(0)}, queries.options().dtype(at::kInt));
        check_rc(cubvh_hashtable_search(h_, queries.device().index(), queries.data_ptr<int32_t>(), queries.size(0),
                                        (int32_t)queries.size(1), out.data_ptr<int32_t>(), current_stream(queries)),
                 "search");
        return out;
    }

private:
    // api_gpu.cu:240-242, 252-254, 264-266
    static at::Tensor checked_keys(const at::Tensor& t, const char* name) {
        TORCH_CHECK(t.is_cuda(), name, " must reside on CUDA");
        TORCH_CHECK(t.scalar_type() == at::kInt, name, " must be int32");
        TORCH_CHECK(t.dim() == 2, name, " must be 2D [N,D]");
        return t.contiguous();
    }

    cubvh_hashtable* h_ = nullptr;
};

struct CpuHashTable {};

}  // namespace

PYBIND11_MODULE(_cubvh, m) {
    m.doc() = "MI355X (gfx950) drop-in for the reference package's native module _cubvh, over libcubvh_amd.so";

    py::class_<BvhImpl>(m, "cuBVH")
        .def("ray_trace", &BvhImpl::ray_trace)
        .def("unsigned_distance", &BvhImpl::unsigned_distance)
        .def("signed_distance", &BvhImpl::signed_distance)
        .def("state_dict", &BvhImpl::state_dict);

    m.def("create_cuBVH", &create_bvh, py::return_value_policy::take_ownership);
    m.def("from_state_dict", &from_state_dict, py::return_value_policy::take_ownership);
    m.def("floodfill", &floodfill);
    m.def("sparse_marching_cubes", &sparse_marching_cubes, py::arg("coords"), py::arg("corners"), py::arg("iso"),
          py::arg("ensure_consistency") = false);

    py::class_<HashTableImpl>(m, "cuHashTable")
        .def("set_num_dims", &HashTableImpl::set_num_dims)
        .def("get_num_dims", &HashTableImpl::get_num_dims)
        .def("resize", &HashTableImpl::resize)
        .def("prepare", &HashTableImpl::prepare)
        .def("insert", &HashTableImpl::insert, py::arg("coords"))
        .def("build", &HashTableImpl::build, py::arg("coords"))
        .def("search", &HashTableImpl::search, py::arg("queries"));
    m.def("create_cuHashTable", []() { return new HashTableImpl(); }, py::return_value_policy::take_ownership);

    // CPU-only names of the reference module (bindings.cpp:46-87): outside the BVH query path.
    for (const char* name : {"fill_holes", "merge_vertices", "sparse_marching_cubes_cpu", "decimate", "parallel_decimate"}) {
        std::string n(name);
        m.def(name, [n](py::args, py::kwargs) { not_implemented(n); });
    }
    py::class_<CpuHashTable>(m, "HashTable").def(py::init([]() -> CpuHashTable* { not_implemented("HashTable"); }));
}
