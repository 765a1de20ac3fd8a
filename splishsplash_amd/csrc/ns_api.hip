// ns_api.hip -- sph_ns_* : the CompactNSearch::NeighborhoodSearch surface the
// reference calls (SPlisHSPlasH/NeighborhoodSearch.h:8-14, call sites in
// Simulation.cpp:145-151, 606-642, 696-756, FluidModel.cpp:216-360), on the
// device grid of grid.hpp. Positions stay caller-owned host memory (the
// pointer is kept, as CompactNSearch does); neighbour lists are produced on
// the device and mirrored to host CSR arrays per (set, set) pair.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "grid.hpp"

namespace sph {

struct PointSet {
    const void *x = nullptr;
    size_t n = 0;
    int dtype = SPH_F32;
    bool dynamic = true;
    void *user = nullptr;
    std::vector<unsigned> zperm;  // last z_sort permutation: new slot -> old index
    // device copies
    DBuf<float4> dx, dsorted;
    DBuf<unsigned> perm_copy, start_copy;
    CellSorter sorter;
    bool uploaded = false;
    double lo[3], hi[3];
    // host CSR per target set
    std::vector<std::vector<unsigned>> off, idx;
};

struct NSHandle {
    double radius = 0;
    bool erase_empty = false;
    std::vector<std::unique_ptr<PointSet>> sets;
    std::vector<std::vector<unsigned char>> active;  // active[i][j]: set i searches set j
    hipStream_t stream = nullptr;
    KernelTimer timer;
    bool have_results = false;
    ~NSHandle() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static void upload(NSHandle &h, PointSet &ps) {
    std::vector<float4> hx(ps.n);
    for (int c = 0; c < 3; c++) {
        ps.lo[c] = 1e300;
        ps.hi[c] = -1e300;
    }
    for (size_t i = 0; i < ps.n; i++) {
        float v[3];
        for (int c = 0; c < 3; c++) {
            v[c] = ps.dtype == SPH_F64 ? (float)static_cast<const double *>(ps.x)[3 * i + c]
                                       : static_cast<const float *>(ps.x)[3 * i + c];
            ps.lo[c] = std::min(ps.lo[c], (double)v[c]);
            ps.hi[c] = std::max(ps.hi[c], (double)v[c]);
        }
        hx[i] = make_float4(v[0], v[1], v[2], 0.f);
    }
    ps.dx.reserve(std::max<size_t>(ps.n, 1));
    if (ps.n) SPH_HIP(hipMemcpyAsync(ps.dx.get(), hx.data(), ps.n * sizeof(float4), hipMemcpyHostToDevice, h.stream));
    SPH_HIP(hipStreamSynchronize(h.stream));
    ps.uploaded = true;
}

static void find(NSHandle &h, bool points_changed) {
    require_device();
    const size_t ns = h.sets.size();
    for (auto &ps : h.sets)
        if (!ps->uploaded || (points_changed && ps->dynamic)) upload(h, *ps);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    bool any = false;
    for (auto &ps : h.sets)
        if (ps->n)
            for (int c = 0; c < 3; c++) {
                lo[c] = std::min(lo[c], ps->lo[c]);
                hi[c] = std::max(hi[c], ps->hi[c]);
                any = true;
            }
    if (!any)
        for (int c = 0; c < 3; c++) lo[c] = hi[c] = 0.0;
    const GridDesc g = make_grid(lo, hi, h.radius);
    const float r2 = (float)h.radius * (float)h.radius;
    // sort every set that some set searches in
    std::vector<bool> needed(ns, false);
    for (size_t i = 0; i < ns; i++)
        for (size_t j = 0; j < ns; j++)
            if (h.active[i][j]) needed[j] = true;
    for (size_t j = 0; j < ns; j++) {
        PointSet &ps = *h.sets[j];
        if (!needed[j] || ps.n == 0) continue;
        ps.sorter.sort(h.stream, ps.dx.get(), (unsigned)ps.n, g, h.timer);
        ps.dsorted.reserve(ps.n);
        permute_f4(h.stream, ps.sorter.perm.get(), (unsigned)ps.n, ps.dx.get(), ps.dsorted.get());
    }
    DBuf<unsigned> cnt, off, out;
    for (size_t i = 0; i < ns; i++) {
        PointSet &qi = *h.sets[i];
        qi.off.assign(ns, std::vector<unsigned>());
        qi.idx.assign(ns, std::vector<unsigned>());
        for (size_t j = 0; j < ns; j++) {
            qi.off[j].assign(qi.n + 1, 0u);
            if (!h.active[i][j] || qi.n == 0 || h.sets[j]->n == 0) continue;
            PointSet &sj = *h.sets[j];
            cnt.reserve(qi.n);
            build_lists_csr(h.stream, qi.dx.get(), (unsigned)qi.n, sj.dsorted.get(), sj.sorter.perm.get(),
                            sj.sorter.start.get(), g, r2, i == j, nullptr, nullptr, cnt.get());
            off.reserve(qi.n + 1);
            exclusive_scan(h.stream, cnt.get(), off.get(), (unsigned)qi.n, qi.sorter.bsum, h.timer);
            SPH_HIP(hipMemcpyAsync(qi.off[j].data(), off.get(), (qi.n + 1) * sizeof(unsigned), hipMemcpyDeviceToHost,
                                   h.stream));
            SPH_HIP(hipStreamSynchronize(h.stream));
            const size_t total = qi.off[j][qi.n];
            qi.idx[j].resize(total);
            if (!total) continue;
            out.reserve(total);
            build_lists_csr(h.stream, qi.dx.get(), (unsigned)qi.n, sj.dsorted.get(), sj.sorter.perm.get(),
                            sj.sorter.start.get(), g, r2, i == j, off.get(), out.get(), nullptr);
            SPH_HIP(hipMemcpyAsync(qi.idx[j].data(), out.get(), total * sizeof(unsigned), hipMemcpyDeviceToHost,
                                   h.stream));
            SPH_HIP(hipStreamSynchronize(h.stream));
        }
    }
    h.have_results = true;
}

// Morton code of 21-bit cell coordinates.
static uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// z_sort: per dynamic set, the order of the Morton codes of the points'
// cells (edge = radius), ties by index.
static void zsort(NSHandle &h) {
    for (auto &psp : h.sets) {
        PointSet &ps = *psp;
        if (!ps.dynamic) {
            ps.zperm.clear();
            continue;
        }
        std::vector<uint64_t> key(ps.n);
        double lo[3] = {1e300, 1e300, 1e300};
        auto X = [&](size_t i, int c) {
            return ps.dtype == SPH_F64 ? static_cast<const double *>(ps.x)[3 * i + c]
                                       : (double)static_cast<const float *>(ps.x)[3 * i + c];
        };
        for (size_t i = 0; i < ps.n; i++)
            for (int c = 0; c < 3; c++) lo[c] = std::min(lo[c], X(i, c));
        const double inv = 1.0 / h.radius;
        for (size_t i = 0; i < ps.n; i++) {
            uint64_t m = 0;
            for (int c = 0; c < 3; c++) {
                const double q = std::floor((X(i, c) - lo[c]) * inv);
                m |= spread3((uint64_t)std::min(std::max(q, 0.0), 2097151.0)) << c;
            }
            key[i] = m;
        }
        ps.zperm.resize(ps.n);
        std::iota(ps.zperm.begin(), ps.zperm.end(), 0u);
        std::stable_sort(ps.zperm.begin(), ps.zperm.end(), [&](unsigned a, unsigned b) { return key[a] < key[b]; });
    }
    h.have_results = false;
}

}  // namespace sph

using namespace sph;

struct sph_ns {
    NSHandle h;
};

#define NSCHK(ns)                                                   \
    SPH_REQUIRE((ns) != nullptr, SPH_ERR_INVALID, "null handle");

extern "C" {

int sph_ns_create(double radius, int erase_empty_cells, sph_ns_t *out) {
    return guarded([&] {
        SPH_REQUIRE(out && radius > 0, SPH_ERR_INVALID, "sph_ns_create: bad arguments");
        require_device();
        auto *ns = new sph_ns();
        ns->h.radius = radius;
        ns->h.erase_empty = erase_empty_cells != 0;
        hipError_t e = hipStreamCreateWithFlags(&ns->h.stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ns;
            throw Error(SPH_ERR_HIP, "hipStreamCreate failed");
        }
        *out = ns;
    });
}

int sph_ns_destroy(sph_ns_t ns) {
    return guarded([&] { delete ns; });
}

int sph_ns_set_radius(sph_ns_t ns, double radius) {
    return guarded([&] {
        NSCHK(ns);
        SPH_REQUIRE(radius > 0, SPH_ERR_INVALID, "radius must be > 0");
        ns->h.radius = radius;
        ns->h.have_results = false;
    });
}

int sph_ns_get_radius(sph_ns_t ns, double *out) {
    return guarded([&] {
        NSCHK(ns);
        *out = ns->h.radius;
    });
}

int sph_ns_reset(sph_ns_t ns) {
    return guarded([&] {
        NSCHK(ns);
        ns->h.sets.clear();
        ns->h.active.clear();
        ns->h.have_results = false;
    });
}

int sph_ns_add_point_set(sph_ns_t ns, const void *x, size_t n, int dtype, int is_dynamic, int search_neighbors,
                         int find_neighbors, void *user_data, unsigned *out_index) {
    return guarded([&] {
        NSCHK(ns);
        SPH_REQUIRE(x || n == 0, SPH_ERR_INVALID, "null positions");
        SPH_REQUIRE(dtype == SPH_F32 || dtype == SPH_F64, SPH_ERR_INVALID, "dtype must be SPH_F32 or SPH_F64");
        auto ps = std::make_unique<PointSet>();
        ps->x = x;
        ps->n = n;
        ps->dtype = dtype;
        ps->dynamic = is_dynamic != 0;
        ps->user = user_data;
        NSHandle &h = ns->h;
        const size_t k = h.sets.size();
        // ActivationTable::add_point_set: row k searches the others when
        // search_neighbors, column k is found by the others when find_neighbors.
        for (auto &row : h.active) row.push_back((unsigned char)(find_neighbors != 0));
        h.active.push_back(std::vector<unsigned char>(k + 1, (unsigned char)(search_neighbors != 0)));
        h.active[k][k] = (unsigned char)(search_neighbors != 0 && find_neighbors != 0);
        h.sets.push_back(std::move(ps));
        h.have_results = false;
        if (out_index) *out_index = (unsigned)k;
    });
}

int sph_ns_resize_point_set(sph_ns_t ns, unsigned set, const void *x, size_t n) {
    return guarded([&] {
        NSCHK(ns);
        SPH_REQUIRE(set < ns->h.sets.size(), SPH_ERR_INVALID, "set index out of range");
        PointSet &ps = *ns->h.sets[set];
        ps.x = x;
        ps.n = n;
        ps.uploaded = false;
        ps.zperm.clear();
        ns->h.have_results = false;
    });
}

int sph_ns_update_point_sets(sph_ns_t ns) {
    return guarded([&] {
        NSCHK(ns);
        require_device();
        for (auto &ps : ns->h.sets)
            if (ps->dynamic || !ps->uploaded) upload(ns->h, *ps);
    });
}

int sph_ns_find_neighbors(sph_ns_t ns, int points_changed) {
    return guarded([&] {
        NSCHK(ns);
        find(ns->h, points_changed != 0);
    });
}

int sph_ns_query_point(sph_ns_t ns, const double x[3], unsigned set, unsigned *out, unsigned cap, unsigned *count) {
    return guarded([&] {
        NSCHK(ns);
        SPH_REQUIRE(set < ns->h.sets.size() && x && count, SPH_ERR_INVALID, "bad arguments");
        PointSet &ps = *ns->h.sets[set];
        if (!ps.uploaded) upload(ns->h, ps);
        // a single point: the host reads the uploaded copy
        std::vector<float4> hx(ps.n);
        if (ps.n)
            SPH_HIP(hipMemcpy(hx.data(), ps.dx.get(), ps.n * sizeof(float4), hipMemcpyDeviceToHost));
        const float q[3] = {(float)x[0], (float)x[1], (float)x[2]};
        const float r2 = (float)ns->h.radius * (float)ns->h.radius;
        unsigned k = 0;
        for (size_t j = 0; j < ps.n; j++) {
            const float dx = q[0] - hx[j].x, dy = q[1] - hx[j].y, dz = q[2] - hx[j].z;
            if ((dx * dx + dy * dy) + dz * dz < r2) {
                if (out && k < cap) out[k] = (unsigned)j;
                k++;
            }
        }
        *count = k;
    });
}

int sph_ns_z_sort(sph_ns_t ns) {
    return guarded([&] {
        NSCHK(ns);
        zsort(ns->h);
    });
}

int sph_ns_sort_field(sph_ns_t ns, unsigned set, void *data, size_t elem_bytes) {
    return guarded([&] {
        NSCHK(ns);
        SPH_REQUIRE(set < ns->h.sets.size() && data && elem_bytes > 0, SPH_ERR_INVALID, "bad arguments");
        PointSet &ps = *ns->h.sets[set];
        if (ps.zperm.empty()) return;
        SPH_REQUIRE(ps.zperm.size() == ps.n, SPH_ERR_STATE, "z_sort permutation does not match the set size");
        std::vector<char> tmp(ps.n * elem_bytes);
        const char *src = static_cast<const char *>(data);
        for (size_t k = 0; k < ps.n; k++) std::memcpy(&tmp[k * elem_bytes], src + (size_t)ps.zperm[k] * elem_bytes, elem_bytes);
        std::memcpy(data, tmp.data(), tmp.size());
    });
}

int sph_ns_set_active_all(sph_ns_t ns, int active) {
    return guarded([&] {
        NSCHK(ns);
        for (auto &row : ns->h.active) std::fill(row.begin(), row.end(), (unsigned char)(active != 0));
    });
}

int sph_ns_set_active_set(sph_ns_t ns, unsigned set, int search_neighbors, int find_neighbors) {
    return guarded([&] {
        NSCHK(ns);
        auto &A = ns->h.active;
        SPH_REQUIRE(set < A.size(), SPH_ERR_INVALID, "set index out of range");
        for (size_t i = 0; i < A.size(); i++) {
            A[i][set] = (unsigned char)(find_neighbors != 0);
            A[set][i] = (unsigned char)(search_neighbors != 0);
        }
        A[set][set] = (unsigned char)(search_neighbors != 0 && find_neighbors != 0);
    });
}

int sph_ns_set_active_pair(sph_ns_t ns, unsigned i, unsigned j, int active) {
    return guarded([&] {
        NSCHK(ns);
        auto &A = ns->h.active;
        SPH_REQUIRE(i < A.size() && j < A.size(), SPH_ERR_INVALID, "set index out of range");
        A[i][j] = (unsigned char)(active != 0);
    });
}

int sph_ns_is_active_pair(sph_ns_t ns, unsigned i, unsigned j, int *out) {
    return guarded([&] {
        NSCHK(ns);
        auto &A = ns->h.active;
        SPH_REQUIRE(i < A.size() && j < A.size() && out, SPH_ERR_INVALID, "bad arguments");
        *out = A[i][j];
    });
}

int sph_ns_n_point_sets(sph_ns_t ns, unsigned *out) {
    return guarded([&] {
        NSCHK(ns);
        *out = (unsigned)ns->h.sets.size();
    });
}

int sph_ns_n_points(sph_ns_t ns, unsigned set, size_t *out) {
    return guarded([&] {
        NSCHK(ns);
        SPH_REQUIRE(set < ns->h.sets.size() && out, SPH_ERR_INVALID, "bad arguments");
        *out = ns->h.sets[set]->n;
    });
}

int sph_ns_user_data(sph_ns_t ns, unsigned set, void **out) {
    return guarded([&] {
        NSCHK(ns);
        SPH_REQUIRE(set < ns->h.sets.size() && out, SPH_ERR_INVALID, "bad arguments");
        *out = ns->h.sets[set]->user;
    });
}

static const PointSet &result_set(sph_ns_t ns, unsigned set_i, unsigned set_j) {
    SPH_REQUIRE(ns != nullptr, SPH_ERR_INVALID, "null handle");
    SPH_REQUIRE(ns->h.have_results, SPH_ERR_STATE, "find_neighbors has not run since the last change");
    SPH_REQUIRE(set_i < ns->h.sets.size() && set_j < ns->h.sets.size(), SPH_ERR_INVALID, "set index out of range");
    return *ns->h.sets[set_i];
}

int sph_ns_n_neighbors(sph_ns_t ns, unsigned set_i, unsigned set_j, unsigned i, unsigned *out) {
    return guarded([&] {
        const PointSet &ps = result_set(ns, set_i, set_j);
        SPH_REQUIRE(i < ps.n && out, SPH_ERR_INVALID, "point index out of range");
        *out = ps.off[set_j][i + 1] - ps.off[set_j][i];
    });
}

int sph_ns_neighbor(sph_ns_t ns, unsigned set_i, unsigned set_j, unsigned i, unsigned k, unsigned *out) {
    return guarded([&] {
        const PointSet &ps = result_set(ns, set_i, set_j);
        SPH_REQUIRE(i < ps.n && out, SPH_ERR_INVALID, "point index out of range");
        const unsigned o = ps.off[set_j][i];
        SPH_REQUIRE(k < ps.off[set_j][i + 1] - o, SPH_ERR_INVALID, "neighbour index out of range");
        *out = ps.idx[set_j][o + k];
    });
}

int sph_ns_neighbor_list(sph_ns_t ns, unsigned set_i, unsigned set_j, unsigned i, const unsigned **out,
                         unsigned *count) {
    return guarded([&] {
        const PointSet &ps = result_set(ns, set_i, set_j);
        SPH_REQUIRE(i < ps.n && out && count, SPH_ERR_INVALID, "bad arguments");
        const unsigned o = ps.off[set_j][i];
        *count = ps.off[set_j][i + 1] - o;
        *out = ps.idx[set_j].data() + o;
    });
}

int sph_ns_neighbor_csr(sph_ns_t ns, unsigned set_i, unsigned set_j, const unsigned **offsets,
                        const unsigned **indices) {
    return guarded([&] {
        const PointSet &ps = result_set(ns, set_i, set_j);
        SPH_REQUIRE(offsets && indices, SPH_ERR_INVALID, "bad arguments");
        *offsets = ps.off[set_j].data();
        *indices = ps.idx[set_j].data();
    });
}

}  // extern "C"
