// C++ caller of the drop-in classes of include/sph_amd.hpp: the calls the
// reference makes on its NeighborhoodSearch (FluidModel/BoundaryModel
// registration, find_neighbors, z_sort + sort_field, set_active, neighbour
// accessors) and a few TimeStepDFSPH steps. Neighbour sets are checked
// against a brute-force FP32 search in this program. Prints "OK" on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "sph_amd.hpp"

static unsigned long long g_seed = 12345;
static float frand() {
    g_seed = g_seed * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((g_seed >> 40) & 0xffffff) / 16777216.0f;
}

#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            return 1;                                                    \
        }                                                                \
    } while (0)

int main() {
    const float r = 0.1f;
    std::vector<float> fluid(3 * 4000), wall;
    for (auto &v : fluid) v = 0.2f + 0.6f * frand();
    for (int i = 0; i <= 20; i++)
        for (int j = 0; j <= 20; j++) {
            wall.push_back(0.05f * i);
            wall.push_back(0.0f);
            wall.push_back(0.05f * j);
        }
    sph_amd::NeighborhoodSearch ns(r, false);
    const unsigned f = ns.add_point_set(fluid.data(), fluid.size() / 3, true, true, true);
    const unsigned b = ns.add_point_set(wall.data(), wall.size() / 3, false, false, true);
    ns.find_neighbors();
    const auto &ps = ns.point_set(f);
    const size_t n = ps.n_points();
    EXPECT(n == 4000 && ns.n_point_sets() == 2);
    const float r2 = r * r;
    for (size_t i = 0; i < n; i += 97) {
        std::set<unsigned> ref, got;
        for (size_t j = 0; j < n; j++) {
            if (j == i) continue;
            const float dx = fluid[3 * i] - fluid[3 * j], dy = fluid[3 * i + 1] - fluid[3 * j + 1],
                        dz = fluid[3 * i + 2] - fluid[3 * j + 2];
            if ((dx * dx + dy * dy) + dz * dz < r2) ref.insert((unsigned)j);
        }
        const unsigned *lst = ps.neighbor_list(f, (unsigned)i);
        for (unsigned k = 0; k < ps.n_neighbors(f, (unsigned)i); k++) got.insert(lst[k]);
        EXPECT(ref == got);
    }
    // z_sort + sort_field keeps positions and ids consistent
    std::vector<unsigned> ids(n);
    for (size_t i = 0; i < n; i++) ids[i] = (unsigned)i;
    const std::vector<float> before = fluid;
    ns.z_sort();
    ps.sort_field(reinterpret_cast<float(*)[3]>(fluid.data()));
    ps.sort_field(ids.data());
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) EXPECT(fluid[3 * i + c] == before[3 * ids[i] + c]);
    ns.find_neighbors();
    // Akinci-style activation: boundary searches only itself
    ns.set_active(false);
    ns.set_active(b, true, true);
    EXPECT(ns.is_active(b, b) && ns.is_active(b, f) && !ns.is_active(f, f));
    ns.find_neighbors();
    EXPECT(ns.point_set(b).n_neighbors(b, 0) > 0);
    // point query
    std::vector<std::vector<unsigned>> q;
    const double x[3] = {0.5, 0.5, 0.5};
    ns.find_neighbors(x, q);
    EXPECT(q.size() == 2);

    // a few DFSPH steps of a small dam break with a volume-map wall
    sph_dfsph_params_t p = sph_amd::TimeStepDFSPH::default_params();
    p.cfl_factor = 1.0;
    p.time_step_size = 0.004;
    std::vector<float> xf, vf;
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++)
            for (int k = 0; k < 12; k++) {
                xf.push_back(0.1f + 0.05f * i);
                xf.push_back(0.1f + 0.05f * j);
                xf.push_back(0.1f + 0.05f * k);
            }
    vf.assign(xf.size(), 0.0f);
    sph_amd::TimeStepDFSPH dfsph(p);
    dfsph.set_fluid(xf.data(), vf.data(), xf.size() / 3);
    const double lo[3] = {0, 0, 0}, hi[3] = {2, 1.5, 1.5};
    const unsigned res[3] = {72, 62, 62};
    dfsph.set_volume_map_box(lo, hi, res);
    dfsph.step(10);
    const auto c = dfsph.counters();
    EXPECT(c.steps == 10 && c.iterations >= 2 && c.flags == 0);
    const auto pos = dfsph.field("position", 3);
    for (float v : pos) EXPECT(std::isfinite(v));
    std::printf("OK\n");
    return 0;
}
