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
// measurements_<id>.txt of save_additional_info, and apbf_scene_save_svg the
// planar drawing particles_<id>.svg of save_as_svg (save_particle_info.cpp:
// 130-270; the reference draws only in 2D builds, :183). The text dumps
// format numbers with std::ostream's defaults and the SVG with std::format's
// "{}" (the shortest round-trip form, std::to_chars), as the reference does.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "scene.h"

namespace {

struct FluidDump {
    std::vector<int32_t> pos;  // 4 per particle
    std::vector<float> radius, kw, trad, inv_mass, boundariness;
    std::vector<uint32_t> bd, pairs, t_src, t_tgt;
    uint32_t n = 0, n_pairs = 0;
};

apbf_status read_fluid(void* scene, FluidDump& d) {
    uint32_t nt = 0;
    apbf_status st = apbf_scene_get_counts(scene, &d.n, &nt, &d.n_pairs);
    if (st != APBF_OK) return st;
    const size_t n = d.n;
    d.pos.resize(4 * n);
    std::vector<float> vel(4 * n), ttl(nt);
    std::vector<int32_t> backup(4 * n);
    std::vector<uint32_t> tr(n);
    d.boundariness.resize(n);
    d.t_src.resize(nt);
    d.t_tgt.resize(nt);
    d.radius.resize(n);
    d.kw.resize(n);
    d.trad.resize(n);
    d.inv_mass.resize(n);
    d.bd.resize(n);
    d.pairs.resize(2 * (size_t)d.n_pairs);
    float ldt = 0.0f;
    return apbf_scene_get_state(scene, d.pos.data(), vel.data(), d.inv_mass.data(), d.radius.data(), backup.data(),
                                tr.data(), d.trad.data(), d.kw.data(), d.boundariness.data(), d.bd.data(),
                                d.t_src.data(), d.t_tgt.data(), ttl.data(), d.pairs.data(), &ldt);
}

// std::format("{}") of a float or an unsigned integer: std::to_chars without
// a format, the shortest string that reads back to the same value.
template <class T>
std::string fmt(T v) {
    char b[64];
    auto r = std::to_chars(b, b + sizeof b, v);
    return std::string(b, r.ptr);
}
std::string hex2(uint32_t v) {  // "{:02x}"
    static const char* digits = "0123456789abcdef";
    std::string out;
    char b[16];
    int k = 0;
    do {
        b[k++] = digits[v & 15u];
        v >>= 4;
    } while (v);
    while (k < 2) b[k++] = '0';
    while (k) out += b[--k];
    return out;
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

apbf_status apbf_scene_save_svg(void* scene, const char* folder, uint32_t svg_id, const float* view_min2,
                                const float* view_max2, float render_scale, float max_expected_boundary_distance,
                                int32_t color, int32_t render_boxes) {
    if (!scene || !folder || !view_min2 || !view_max2) return APBF_ERR_INVALID_ARGUMENT;
    apbf::Scene* sc = static_cast<apbf::Scene*>(scene);
    if (sc->dims() > 2u) return APBF_OK;  // save_particle_info.cpp:183: no drawing in 3D
    FluidDump d;
    apbf_status st = read_fluid(scene, d);
    if (st != APBF_OK) return st;
    const uint32_t n = d.n;
    const float res = 262144.0f;  // POS_RESOLUTION
    // save_particle_info.cpp:132-142
    const bool include_boxes = sc->boxes.n > 0 && !sc->packed() && render_boxes != 0;
    const bool include_kernels = color == 3;
    bool gradient = false;
    uint32_t color_count = 2;
    if (color == 1) {
        gradient = true;
        color_count = 0;
    } else if (color == 2) {
        color_count = 10;
    }
    const float stroke = std::min(view_max2[0] - view_min2[0], view_max2[1] - view_min2[1]) / 1000.0f;
    // :161-173; the fluid id is the particle's slot here (indices[i] == i)
    std::vector<uint32_t> pcolor(n, 0u);
    if (color == 2) {
        for (size_t i = 0; i < d.t_src.size(); ++i) {
            const uint32_t c = (uint32_t)(i % (color_count - 1u)) + 1u;
            if (d.t_src[i] < n) pcolor[d.t_src[i]] = c;
            if (d.t_tgt[i] < n) pcolor[d.t_tgt[i]] = c;
        }
    } else if (color != 1) {
        for (uint32_t i = 0; i < n; ++i) pcolor[i] = d.boundariness[i] >= 1.0f ? 1u : 0u;
    }
    // :185-199
    static const char* fills[10] = {"0000ff", "ff0000", "00ff00", "ff8000", "ffff00",
                                     "ff00ff", "00ffff", "008000", "800000", "ffaec8"};
    std::string originals, particles, kernels;
    for (uint32_t c = 0; c < color_count; ++c)
        originals += "<circle cx=\"0\" cy=\"0\" r=\"" + fmt(render_scale) + "\" id=\"p" + fmt(c) +
                     "\" style=\"fill:#" + fills[c] + ";stroke-width:1\" />";
    if (include_kernels)
        originals += "<circle cx=\"0\" cy=\"0\" r=\"1\" id=\"k\" style=\"fill:none;stroke:#ff7f00;stroke-width:" +
                     fmt(stroke) + ";vector-effect:non-scaling-stroke\" />";
    if (include_boxes)
        originals += "<rect x=\"0\" y=\"0\" width=\"1\" height=\"1\" id=\"box\" style=\"fill:#bffeff;stroke-width:1\" />";
    // :201-228
    const float c2 = max_expected_boundary_distance * res * 0.8f;
    for (uint32_t i = 0; i < n; ++i) {
        const float px = (float)d.pos[4 * (size_t)i] / res;
        const float py = -((float)d.pos[4 * (size_t)i + 1] / res);
        const float rad = d.radius[i];
        const float ker = d.kw[i];
        if (gradient) {
            float a = ((float)d.bd[i] - 0.0f) / (c2 - 0.0f);
            a = std::min(std::max(a, 0.0f), 1.0f);
            // glm::mix(x, y, a) = x * (1 - a) + y * a, then * 255 and truncation
            const float from[3] = {0.0f, 0.0f, 1.0f}, to[3] = {0.62f, 0.96f, 0.83f};
            uint32_t col[3];
            for (int k = 0; k < 3; ++k) {
                const float m = from[k] * (1.0f - a) + to[k] * a;
                col[k] = (uint32_t)(m * 255.0f);
            }
            particles += "<circle cx=\"" + fmt(px) + "\" cy=\"" + fmt(py) + "\" r=\"" + fmt(rad * render_scale) +
                         "\" style=\"fill:#" + hex2(col[0]) + hex2(col[1]) + hex2(col[2]) + ";stroke-width:1\" />";
        } else {
            particles += "<use transform=\"matrix(" + fmt(rad) + ",0,0," + fmt(rad) + "," + fmt(px) + "," + fmt(py) +
                         ")\" xlink:href=\"#p" + fmt(pcolor[i]) + "\" />";
        }
        if (include_kernels)
            kernels += "<use transform=\"matrix(" + fmt(ker) + ",0,0," + fmt(ker) + "," + fmt(px) + "," + fmt(py) +
                       ")\" xlink:href=\"#k\" />";
    }
    // :230-236, boxes_to_svg :256-269
    std::string svg;
    if (!originals.empty()) svg += "<g id=\"originals\" style=\"display:none\">" + originals + "</g>";
    if (include_boxes) {
        std::string b;
        for (uint32_t i = 0; i < sc->boxes.n; ++i) {
            const float* mn = sc->boxes.mn[i];
            const float* mx = sc->boxes.mx[i];
            b += "<use transform=\"matrix(" + fmt(mx[0] - mn[0]) + ",0,0," + fmt(mx[1] - mn[1]) + "," + fmt(mn[0]) +
                 "," + fmt(-mx[1]) + ")\" xlink:href=\"#box\" />";
        }
        svg += "<g id=\"boxes\">" + b + "</g>";
    }
    if (include_kernels) svg += "<g id=\"kernels\">" + kernels + "</g>";
    svg += "<g id=\"particles\">" + particles + "</g>";
    svg = "<svg viewBox=\"" + fmt(view_min2[0]) + " " + fmt(-view_max2[1]) + " " + fmt(view_max2[0] - view_min2[0]) +
          " " + fmt(view_max2[1] - view_min2[1]) +
          "\" version=\"1.1\" xmlns=\"http://www.w3.org/2000/svg\" xmlns:xlink=\"http://www.w3.org/1999/xlink\" "
          "xmlns:svg=\"http://www.w3.org/2000/svg\">" +
          svg + "</svg>";
    std::error_code ec;
    std::filesystem::create_directories(folder, ec);
    std::ofstream o(std::string(folder) + "/particles_" + std::to_string(svg_id) + ".svg");
    o << svg;
    return o ? APBF_OK : APBF_ERR_INVALID_ARGUMENT;
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
