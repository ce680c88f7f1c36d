// save_particle_info (save_particle_info.cpp:21-128, 271-280): text dumps of a
// scene's fluid, in the reference's file names and formats, so that a state of
// this library can be compared with a dump of a reference run. Host code only:
// the state comes through apbf_scene_get_state.
//
// Files of apbf_scene_save_particle_info(folder) (the reference writes them to
// PARTICLE_INFO_FOLDER_NAME "particle_data", cpu_gpu_shared_config.h:21):
// centerDist.txt, radius.txt, neighborCount.txt, kernelWidth.txt,
// targetRadius.txt, boundaryDistance.txt (every value followed by ';') and
// data.csv (a header line, then one line per particle in ascending distance
// to the centre (0, 10, -60)). apbf_scene_save_measurements writes
// measurements_<id>.txt of save_additional_info; in 3D the reference's SVG
// dump stops there (save_particle_info.cpp:185), and so does this one.
// Numbers go through std::ostream's default formatting, as in the reference.
#include <algorithm>
#include <filesystem>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "scene.h"

namespace {

struct FluidDump {
    std::vector<int32_t> pos;  // 4 per particle
    std::vector<float> radius, kw, trad, inv_mass;
    std::vector<uint32_t> bd, pairs;
    uint32_t n = 0, n_pairs = 0;
};

apbf_status read_fluid(void* scene, FluidDump& d) {
    uint32_t nt = 0;
    apbf_status st = apbf_scene_get_counts(scene, &d.n, &nt, &d.n_pairs);
    if (st != APBF_OK) return st;
    const size_t n = d.n;
    d.pos.resize(4 * n);
    std::vector<float> vel(4 * n), boundariness(n), ttl(nt);
    std::vector<int32_t> backup(4 * n);
    std::vector<uint32_t> tr(n), ts(nt), tt(nt);
    d.radius.resize(n);
    d.kw.resize(n);
    d.trad.resize(n);
    d.inv_mass.resize(n);
    d.bd.resize(n);
    d.pairs.resize(2 * (size_t)d.n_pairs);
    float ldt = 0.0f;
    return apbf_scene_get_state(scene, d.pos.data(), vel.data(), d.inv_mass.data(), d.radius.data(), backup.data(),
                                tr.data(), d.trad.data(), d.kw.data(), boundariness.data(), d.bd.data(), ts.data(),
                                tt.data(), ttl.data(), d.pairs.data(), &ldt);
}

}  // namespace

extern "C" {

apbf_status apbf_scene_save_particle_info(void* scene, const char* folder) {
    if (!scene || !folder) return APBF_ERR_INVALID_ARGUMENT;
    FluidDump d;
    apbf_status st = read_fluid(scene, d);
    if (st != APBF_OK) return st;
    const uint32_t n = d.n;
    const float res = 262144.0f;  // POS_RESOLUTION
    const float cx = 0.0f, cy = 10.0f, cz = -60.0f;  // centerPos, save_particle_info.cpp:45
    std::vector<uint32_t> count(n, 0u);
    for (uint32_t k = 0; k < d.n_pairs; ++k)
        if (d.pairs[2 * (size_t)k] < n) ++count[d.pairs[2 * (size_t)k]];
    std::vector<float> center(n), bdr(n), px(n), py(n), pz(n);
    for (uint32_t i = 0; i < n; ++i) {
        px[i] = (float)d.pos[4 * (size_t)i] / res;
        py[i] = (float)d.pos[4 * (size_t)i + 1] / res;
        pz[i] = (float)d.pos[4 * (size_t)i + 2] / res;
        const float dx = px[i] - cx, dy = py[i] - cy, dz = pz[i] - cz;
        center[i] = std::sqrt((dx * dx + dy * dy) + dz * dz);
        bdr[i] = (float)d.bd[i] / res;
    }
    std::error_code ec;
    std::filesystem::create_directories(folder, ec);
    const std::string f(folder);
    auto write_list = [&](const char* name, auto&& value) {
        std::ofstream o(f + "/" + name);
        for (uint32_t i = 0; i < n; ++i) o << value(i) << ";";
        return (bool)o;
    };
    bool ok = write_list("centerDist.txt", [&](uint32_t i) { return center[i]; }) &&
              write_list("radius.txt", [&](uint32_t i) { return d.radius[i]; }) &&
              write_list("neighborCount.txt", [&](uint32_t i) { return count[i]; }) &&
              write_list("kernelWidth.txt", [&](uint32_t i) { return d.kw[i]; }) &&
              write_list("targetRadius.txt", [&](uint32_t i) { return d.trad[i]; }) &&
              write_list("boundaryDistance.txt", [&](uint32_t i) { return bdr[i]; });
    // data.csv in ascending centre distance (std::sort there; a stable sort here
    // fixes the order of equal distances)
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return center[a] < center[b]; });
    std::ofstream o(f + "/data.csv");
    o << "center distance,boundary distance,kernel width,neighbor count,radius,target radius,inverse mass,x,y,z"
      << std::endl;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = order[k];
        o << center[i] << "," << bdr[i] << "," << d.kw[i] << "," << count[i] << "," << d.radius[i] << ","
          << d.trad[i] << "," << d.inv_mass[i] << "," << px[i] << "," << py[i] << "," << pz[i] << std::endl;
    }
    ok = ok && (bool)o;
    return ok ? APBF_OK : APBF_ERR_INVALID_ARGUMENT;
}

apbf_status apbf_scene_save_measurements(void* scene, const char* folder, uint32_t id, float step_ms,
                                         float neighborhood_ms, float solver_ms) {
    if (!scene || !folder) return APBF_ERR_INVALID_ARGUMENT;
    uint32_t n = 0, nt = 0, np = 0;
    apbf_status st = apbf_scene_get_counts(scene, &n, &nt, &np);
    if (st != APBF_OK) return st;
    static const char* names[3] = {"neighborhood_green.h", "neighborhood_binary_search.h",
                                   "neighborhood_brute_force.h"};
    const int m = static_cast<apbf::Scene*>(scene)->nb_method;
    std::error_code ec;
    std::filesystem::create_directories(folder, ec);
    std::ofstream o(std::string(folder) + "/measurements_" + std::to_string(id) + ".txt");
    o << "Particle Count      : " << n << std::endl;
    o << "Neighbor Pair Count : " << np << std::endl;
    o << "Simulation Step     : " << step_ms << "ms" << std::endl;
    o << "Neighborhood Search : " << neighborhood_ms << "ms" << std::endl;
    o << "Constraint Solver   : " << solver_ms << "ms" << std::endl;
    o << "Neighbor Search Type: " << names[m >= 0 && m < 3 ? m : 0] << std::endl;
    return o ? APBF_OK : APBF_ERR_INVALID_ARGUMENT;
}

}  // extern "C"
