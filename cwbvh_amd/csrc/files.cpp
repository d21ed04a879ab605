// On-disk formats (host).
//
//  * mesh.dat, the reference's scene file (src/main.cpp:31-45): int32 vertex
//    count, int32 triangle count, float3 positions, int3 indices, int32
//    material per triangle; little-endian, packed.
//  * samplecache.dat, the reference's surfel file (src/main.cpp:52-64): int32
//    size x, int32 size y, float4 world positions, float4 world normals,
//    float texel radii (size x * size y of each).
//  * layout cache: a built BVH2 or CWBVH layout, so a scene can be loaded
//    instead of rebuilt.  Header, the three arrays, and a 64-bit FNV-1a hash
//    of everything before it.  Loading goes through the same structural
//    validation as cwbvh_layout_from_arrays.
//
// Every reader checks that the file holds exactly the bytes its header
// announces before it allocates, so a truncated or foreign file is refused
// with CWBVH_ERR_IO / CWBVH_ERR_INVALID_ARGUMENT rather than read past.
#include <cstdio>
#include <memory>

#include "common.h"

using namespace cw;

struct cwbvh_samples {
    int32_t size_x = 0, size_y = 0;
    std::vector<cwbvh_float4> positions, normals;
    std::vector<float> radius;
};

namespace {

struct FileCloser {
    void operator()(std::FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<std::FILE, FileCloser>;

int32_t io_error(const std::string& msg) {
    set_error(msg);
    return CWBVH_ERR_IO;
}

int64_t file_size(std::FILE* f) {
    if (std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long n = std::ftell(f);
    if (std::fseek(f, 0, SEEK_SET) != 0) return -1;
    return (int64_t)n;
}

bool read_all(std::FILE* f, void* dst, size_t bytes) { return bytes == 0 || std::fread(dst, 1, bytes, f) == bytes; }
bool write_all(std::FILE* f, const void* src, size_t bytes) {
    return bytes == 0 || std::fwrite(src, 1, bytes, f) == bytes;
}

// 64-bit FNV-1a, streamed.
struct Fnv {
    uint64_t h = 0xcbf29ce484222325ULL;
    void add(const void* p, size_t n) {
        const auto* b = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 0x100000001b3ULL;
        }
    }
};

constexpr char kLayoutMagic[8] = {'C', 'W', 'B', 'V', 'H', 'L', 'Y', '1'};
constexpr uint32_t kLayoutVersion = 1;

struct LayoutHeader {
    char magic[8];
    uint32_t version;
    int32_t kind;
    int32_t node_stride;
    int32_t reserved;
    int64_t nodes_f4;
    int64_t woop_f4;
    int64_t prim_idx_len;
    double sah_cost;
};
static_assert(sizeof(LayoutHeader) == 56, "packed header");

}  // namespace

extern "C" {

int32_t cwbvh_mesh_read_dat(const char* path, cwbvh_mesh** out) {
    CW_CHECK_ARG(path != nullptr && out != nullptr, "NULL argument");
    File f(std::fopen(path, "rb"));
    if (!f) return io_error(std::string("cannot open ") + path);
    const int64_t size = file_size(f.get());
    int32_t nv = 0, nt = 0;
    if (!read_all(f.get(), &nv, 4) || !read_all(f.get(), &nt, 4)) return io_error("mesh file shorter than its header");
    CW_CHECK_ARG(nv >= 0 && nt >= 0, "negative count in the mesh header");
    const int64_t want = 8 + 12 * (int64_t)nv + 12 * (int64_t)nt + 4 * (int64_t)nt;
    if (size != want)
        return io_error("mesh file is " + std::to_string(size) + " bytes, its header announces " + std::to_string(want));
    std::unique_ptr<cwbvh_mesh> m(new cwbvh_mesh());
    m->vertices.resize(3 * (size_t)nv);
    m->indices.resize(3 * (size_t)nt);
    m->materials.resize((size_t)nt);
    if (!read_all(f.get(), m->vertices.data(), 12 * (size_t)nv) ||
        !read_all(f.get(), m->indices.data(), 12 * (size_t)nt) ||
        !read_all(f.get(), m->materials.data(), 4 * (size_t)nt))
        return io_error("short read of the mesh file");
    for (int32_t v : m->indices) CW_CHECK_ARG(v >= 0 && v < nv, "vertex index out of range in the mesh file");
    *out = m.release();
    return CWBVH_OK;
}

int32_t cwbvh_mesh_write_dat(const cwbvh_mesh* mesh, const int32_t* materials, const char* path) {
    CW_CHECK_ARG(mesh != nullptr && path != nullptr, "NULL argument");
    CW_CHECK_ARG(mesh->nv() <= INT32_MAX && mesh->nt() <= INT32_MAX, "mesh too large for the int32 header");
    const int32_t nv = (int32_t)mesh->nv(), nt = (int32_t)mesh->nt();
    std::vector<int32_t> mat;
    if (materials) mat.assign(materials, materials + nt);
    else if ((int64_t)mesh->materials.size() == nt) mat = mesh->materials;
    else mat.assign((size_t)nt, 0);
    File f(std::fopen(path, "wb"));
    if (!f) return io_error(std::string("cannot create ") + path);
    if (!write_all(f.get(), &nv, 4) || !write_all(f.get(), &nt, 4) ||
        !write_all(f.get(), mesh->vertices.data(), 12 * (size_t)nv) ||
        !write_all(f.get(), mesh->indices.data(), 12 * (size_t)nt) || !write_all(f.get(), mat.data(), 4 * (size_t)nt))
        return io_error(std::string("short write to ") + path);
    if (std::fflush(f.get()) != 0) return io_error(std::string("cannot flush ") + path);
    return CWBVH_OK;
}

int32_t cwbvh_mesh_materials(const cwbvh_mesh* mesh, const int32_t** materials, int64_t* count) {
    CW_CHECK_ARG(mesh != nullptr && materials != nullptr && count != nullptr, "NULL argument");
    *materials = mesh->materials.empty() ? nullptr : mesh->materials.data();
    *count = (int64_t)mesh->materials.size();
    return CWBVH_OK;
}

int32_t cwbvh_samples_read(const char* path, cwbvh_samples** out) {
    CW_CHECK_ARG(path != nullptr && out != nullptr, "NULL argument");
    File f(std::fopen(path, "rb"));
    if (!f) return io_error(std::string("cannot open ") + path);
    const int64_t size = file_size(f.get());
    int32_t sx = 0, sy = 0;
    if (!read_all(f.get(), &sx, 4) || !read_all(f.get(), &sy, 4)) return io_error("sample file shorter than its header");
    CW_CHECK_ARG(sx >= 0 && sy >= 0, "negative size in the sample header");
    const int64_t n = (int64_t)sx * sy;
    const int64_t want = 8 + n * (16 + 16 + 4);
    if (size != want)
        return io_error("sample file is " + std::to_string(size) + " bytes, its header announces " +
                        std::to_string(want));
    std::unique_ptr<cwbvh_samples> s(new cwbvh_samples());
    s->size_x = sx;
    s->size_y = sy;
    s->positions.resize((size_t)n);
    s->normals.resize((size_t)n);
    s->radius.resize((size_t)n);
    if (!read_all(f.get(), s->positions.data(), 16 * (size_t)n) ||
        !read_all(f.get(), s->normals.data(), 16 * (size_t)n) || !read_all(f.get(), s->radius.data(), 4 * (size_t)n))
        return io_error("short read of the sample file");
    *out = s.release();
    return CWBVH_OK;
}

int32_t cwbvh_samples_write(const char* path, int32_t size_x, int32_t size_y, const cwbvh_float4* positions,
                            const cwbvh_float4* normals, const float* radius) {
    CW_CHECK_ARG(path != nullptr, "NULL argument");
    CW_CHECK_ARG(size_x >= 0 && size_y >= 0, "negative size");
    const int64_t n = (int64_t)size_x * size_y;
    CW_CHECK_ARG(n == 0 || (positions && normals && radius), "NULL sample array");
    File f(std::fopen(path, "wb"));
    if (!f) return io_error(std::string("cannot create ") + path);
    if (!write_all(f.get(), &size_x, 4) || !write_all(f.get(), &size_y, 4) ||
        !write_all(f.get(), positions, 16 * (size_t)n) || !write_all(f.get(), normals, 16 * (size_t)n) ||
        !write_all(f.get(), radius, 4 * (size_t)n))
        return io_error(std::string("short write to ") + path);
    if (std::fflush(f.get()) != 0) return io_error(std::string("cannot flush ") + path);
    return CWBVH_OK;
}

int32_t cwbvh_samples_view(const cwbvh_samples* s, int32_t* size_x, int32_t* size_y, const cwbvh_float4** positions,
                           const cwbvh_float4** normals, const float** radius) {
    CW_CHECK_ARG(s != nullptr, "samples is NULL");
    if (size_x) *size_x = s->size_x;
    if (size_y) *size_y = s->size_y;
    if (positions) *positions = s->positions.data();
    if (normals) *normals = s->normals.data();
    if (radius) *radius = s->radius.data();
    return CWBVH_OK;
}

void cwbvh_samples_free(cwbvh_samples* s) { delete s; }

int32_t cwbvh_layout_save(const cwbvh_layout* L, const char* path) {
    CW_CHECK_ARG(L != nullptr && path != nullptr, "NULL argument");
    LayoutHeader h{};
    std::memcpy(h.magic, kLayoutMagic, 8);
    h.version = kLayoutVersion;
    h.kind = L->kind;
    h.node_stride = L->node_stride;
    h.nodes_f4 = (int64_t)L->nodes.size();
    h.woop_f4 = (int64_t)L->woop.size();
    h.prim_idx_len = (int64_t)L->prim_idx.size();
    h.sah_cost = L->sah_cost;
    Fnv fnv;
    fnv.add(&h, sizeof h);
    fnv.add(L->nodes.data(), L->nodes.size() * sizeof(cwbvh_float4));
    fnv.add(L->woop.data(), L->woop.size() * sizeof(cwbvh_float4));
    fnv.add(L->prim_idx.data(), L->prim_idx.size() * sizeof(int32_t));
    File f(std::fopen(path, "wb"));
    if (!f) return io_error(std::string("cannot create ") + path);
    if (!write_all(f.get(), &h, sizeof h) ||
        !write_all(f.get(), L->nodes.data(), L->nodes.size() * sizeof(cwbvh_float4)) ||
        !write_all(f.get(), L->woop.data(), L->woop.size() * sizeof(cwbvh_float4)) ||
        !write_all(f.get(), L->prim_idx.data(), L->prim_idx.size() * sizeof(int32_t)) ||
        !write_all(f.get(), &fnv.h, 8))
        return io_error(std::string("short write to ") + path);
    if (std::fflush(f.get()) != 0) return io_error(std::string("cannot flush ") + path);
    return CWBVH_OK;
}

int32_t cwbvh_layout_load(const char* path, cwbvh_layout** out) {
    CW_CHECK_ARG(path != nullptr && out != nullptr, "NULL argument");
    File f(std::fopen(path, "rb"));
    if (!f) return io_error(std::string("cannot open ") + path);
    const int64_t size = file_size(f.get());
    LayoutHeader h{};
    if (!read_all(f.get(), &h, sizeof h)) return io_error("layout file shorter than its header");
    CW_CHECK_ARG(std::memcmp(h.magic, kLayoutMagic, 8) == 0, "not a layout cache file");
    CW_CHECK_ARG(h.version == kLayoutVersion, "unsupported layout cache version");
    CW_CHECK_ARG(h.nodes_f4 >= 0 && h.woop_f4 >= 0 && h.prim_idx_len >= 0 && h.nodes_f4 < (1ll << 40) &&
                     h.woop_f4 < (1ll << 40) && h.prim_idx_len < (1ll << 40),
                 "bad array sizes in the layout header");
    const int64_t want = (int64_t)sizeof h + 16 * h.nodes_f4 + 16 * h.woop_f4 + 4 * h.prim_idx_len + 8;
    if (size != want)
        return io_error("layout file is " + std::to_string(size) + " bytes, its header announces " +
                        std::to_string(want));
    std::vector<cwbvh_float4> nodes((size_t)h.nodes_f4), woop((size_t)h.woop_f4);
    std::vector<int32_t> prim((size_t)h.prim_idx_len);
    uint64_t stored = 0;
    if (!read_all(f.get(), nodes.data(), nodes.size() * sizeof(cwbvh_float4)) ||
        !read_all(f.get(), woop.data(), woop.size() * sizeof(cwbvh_float4)) ||
        !read_all(f.get(), prim.data(), prim.size() * sizeof(int32_t)) || !read_all(f.get(), &stored, 8))
        return io_error("short read of the layout file");
    Fnv fnv;
    fnv.add(&h, sizeof h);
    fnv.add(nodes.data(), nodes.size() * sizeof(cwbvh_float4));
    fnv.add(woop.data(), woop.size() * sizeof(cwbvh_float4));
    fnv.add(prim.data(), prim.size() * sizeof(int32_t));
    CW_CHECK_ARG(fnv.h == stored, "layout cache checksum mismatch");
    cwbvh_layout* L = nullptr;
    const int32_t rc = cwbvh_layout_from_arrays(h.kind, h.node_stride, nodes.data(), h.nodes_f4, woop.data(),
                                                h.woop_f4, prim.data(), h.prim_idx_len, &L);
    if (rc) return rc;
    L->sah_cost = h.sah_cost;
    *out = L;
    return CWBVH_OK;
}

}  // extern "C"
