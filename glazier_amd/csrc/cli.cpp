// cli.cpp - the `glazier` command line over the HIP engine.
//
// The flags, outputs and exit codes of glazier/src/main.rs: `--model
// <blueprint.json> --out <dir> [--engine hip|gpu] [--trace]`, or `--model
// <blueprint.json> --check`. Writes labels_NNNNN.npy per dump (u32, npy.rs
// layout), summary.json with the reference's keys (main.rs:167-185) plus
// throughput, and one stdout line (main.rs:205-209). The serial `cpu` engine
// is the reference's glazier-core and is not part of this engine.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/glazier_hip.h"
#include "blueprint.h"

using namespace ghip;

namespace {

int usage() {
    std::fprintf(stderr,
                 "usage: glazier --model <blueprint.json> --out <dir> [--engine hip|gpu]\n"
                 "       glazier --model <blueprint.json> --check\n"
                 "\n"
                 "Writes labels_NNNNN.npy per dump and summary.json at the end.\n"
                 "--check parses the description and prints what this engine read, so a\n"
                 "second reader of the same file can be compared against it.\n"
                 "--trace writes trace.csv, one row per step: live cells, mean volume,\n"
                 "and the total of every field.\n");
    return 2;
}

std::string dirname_of(const std::string &p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k);
}

std::string join(const std::string &base, const std::string &name) {
    if (!name.empty() && name[0] == '/') return name;
    if (base.empty()) return name;
    return base + "/" + name;
}

bool make_dirs(const std::string &path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::string fmt_num(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    std::string s = buf;
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

std::string pad5(uint64_t v) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "%05llu", (unsigned long long)v);
    return buf;
}

}  // namespace

int main(int argc, char **argv) {
    std::string model_path, out_dir, engine = "hip";
    bool check = false, trace = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--model" && i + 1 < argc) model_path = argv[++i];
        else if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
        else if (a == "--engine" && i + 1 < argc) engine = argv[++i];
        else if (a == "--check") check = true;
        else if (a == "--trace") trace = true;
        else return usage();
    }
    if (model_path.empty()) return usage();
    std::ifstream in(model_path);
    if (!in) {
        std::fprintf(stderr, "cannot read %s\n", model_path.c_str());
        return 1;
    }
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    Blueprint bp;
    std::string err;
    if (!blueprint_from_json(text, &bp, &err)) {
        std::fprintf(stderr, "%s: %s\n", model_path.c_str(), err.c_str());
        return 1;
    }
    if (check) {
        std::printf("%s\n", check_json(bp).c_str());
        return 0;
    }
    if (out_dir.empty()) return usage();
    if (engine != "hip" && engine != "gpu") {
        std::fprintf(stderr, "unknown engine %s (this build runs the hip engine; the serial "
                             "engine is glazier-core's)\n",
                     engine.c_str());
        return usage();
    }
    if (!make_dirs(out_dir)) {
        std::fprintf(stderr, "cannot create %s\n", out_dir.c_str());
        return 1;
    }
    // Blueprint::simulation (blueprint.rs:336-375)
    const std::string base = dirname_of(model_path);
    ghip_model m = bp.model();
    std::vector<double> nematic;
    if (bp.has_nematic_field) {
        // blueprint.rs:347-358: resolved against the description's directory
        const std::string path = bp.nematic_field.size() && bp.nematic_field[0] == '/'
                                     ? bp.nematic_field
                                     : join(base, bp.nematic_field);
        if (!npy_read_pairs(path, &nematic, &err)) {
            std::fprintf(stderr, "%s: %s\n", path.c_str(), err.c_str());
            return 1;
        }
        const size_t sites = (size_t)m.width * m.height * m.depth;
        if (nematic.size() != 2 * sites) {
            std::fprintf(stderr, "%s: holds %zu sites for a lattice of %zu\n", path.c_str(),
                         nematic.size() / 2, sites);
            return 1;
        }
        for (uint32_t t = 0; t < m.n_types; ++t)
            if (m.types[t].lambda_nematic != 0.0) m.has_nematic = 1;
    }
    std::vector<uint32_t> labels;
    size_t cells = 0;
    if (bp.has_labels) {
        if (!npy_read_labels(join(base, bp.labels), &labels, &err)) {
            std::fprintf(stderr, "%s: %s\n", join(base, bp.labels).c_str(), err.c_str());
            return 1;
        }
        if (labels.size() != (size_t)m.width * m.height * m.depth) {
            std::fprintf(stderr, "the label field has %zu sites for a lattice of %zu\n", labels.size(),
                         (size_t)m.width * m.height * m.depth);
            return 1;
        }
        for (uint32_t l : labels) cells = std::max<size_t>(cells, l);
        if (cells == 0) {
            std::fprintf(stderr, "the label field holds no cells\n");
            return 1;
        }
    } else {
        if (m.n_types < 2) {
            std::fprintf(stderr, "a tiled start needs a cell type beside the medium\n");
            return 1;
        }
        cells = tile_grid(m, bp.side, bp.nx, bp.ny, bp.nz, &labels);
    }
    std::vector<uint8_t> types(cells + 1, 1);
    std::vector<uint8_t> assigned = bp.initial_types(cells);
    for (size_t l = 0; l < cells; ++l) types[l + 1] = assigned[l];

    const auto began = std::chrono::steady_clock::now();
    ghip_sim *sim = nullptr;
    ghip_status st = ghip_create_nematic(&m, labels.data(), labels.size(), types.data(), types.size(),
                                         nullptr, nullptr, nematic.empty() ? nullptr : nematic.data(), &sim);
    if (st != GHIP_OK) {
        std::fprintf(stderr, "%s\n", ghip_last_error());
        return 1;
    }
    const double setup =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - began).count();
    const size_t n = labels.size();
    std::string rows;
    if (trace) {
        rows = "step,cells,mean_volume";
        for (const auto &f : bp.fields) rows += "," + f.name;
        rows += "\n";
    }
    std::vector<uint32_t> vol;
    std::vector<double> field(n);
    auto live_stats = [&](size_t *live, double *mean) -> bool {
        size_t nl = 0;
        if (ghip_n_labels(sim, &nl) != GHIP_OK) return false;
        vol.resize(nl);
        if (ghip_volumes(sim, vol.data(), nl) != GHIP_OK) return false;
        size_t c = 0;
        double sum = 0.0;
        for (size_t l = 1; l < nl; ++l)
            if (vol[l] > 0) {
                ++c;
                sum += vol[l];
            }
        *live = c;
        *mean = c ? sum / (double)c : 0.0;
        return true;
    };
    auto field_totals = [&](std::vector<double> *totals) -> bool {
        totals->clear();
        for (uint32_t k = 0; k < m.n_species; ++k) {
            if (ghip_field(sim, k, field.data(), n) != GHIP_OK) return false;
            double t = 0.0;
            for (double v : field) t += v;
            totals->push_back(t);
        }
        return true;
    };
    // step in chunks between dumps (main.rs:297-316); with --trace, one step
    // at a time so every step gets a row
    const uint64_t chunk = trace ? 1 : (bp.dump_every > 0 ? bp.dump_every : bp.steps);
    uint64_t done = 0;
    std::vector<uint32_t> lab(n);
    while (done < bp.steps) {
        const uint64_t take = std::min<uint64_t>(chunk, bp.steps - done);
        if (ghip_step(sim, take) != GHIP_OK) {
            std::fprintf(stderr, "%s\n", ghip_last_error());
            ghip_destroy(sim);
            return 1;
        }
        done += take;
        if (trace) {
            size_t live;
            double mean;
            std::vector<double> totals;
            if (!live_stats(&live, &mean) || !field_totals(&totals)) {
                std::fprintf(stderr, "%s\n", ghip_last_error());
                return 1;
            }
            char buf[64];
            std::snprintf(buf, sizeof buf, "%llu,%zu,%.3f", (unsigned long long)done, live, mean);
            rows += buf;
            for (double t : totals) {
                std::snprintf(buf, sizeof buf, ",%.6f", t);
                rows += buf;
            }
            rows += "\n";
        }
        if (bp.dump_every > 0 && done % bp.dump_every == 0 && done != bp.steps) {
            ghip_labels(sim, lab.data(), n);
            npy_write_u32(join(out_dir, "labels_" + pad5(done) + ".npy"), lab.data(), m.width, m.height,
                          m.depth, &err);
            for (uint32_t k = 0; k < m.n_species; ++k) {
                ghip_field(sim, k, field.data(), n);
                npy_write_f64(join(out_dir, "field_" + bp.fields[k].name + "_" + pad5(done) + ".npy"),
                              field.data(), m.width, m.height, m.depth, &err);
            }
        }
    }
    const double stepping = std::chrono::duration<double>(std::chrono::steady_clock::now() - began).count();
    if (trace) {
        std::ofstream t(join(out_dir, "trace.csv"));
        t << rows;
    }
    std::vector<double> totals;
    size_t live = 0;
    double mean = 0.0;
    if (!field_totals(&totals) || !live_stats(&live, &mean) || ghip_labels(sim, lab.data(), n) != GHIP_OK) {
        std::fprintf(stderr, "%s\n", ghip_last_error());
        return 1;
    }
    uint64_t attempted = 0, accepted = 0;
    ghip_counters(sim, &attempted, &accepted);
    const double overall = std::chrono::duration<double>(std::chrono::steady_clock::now() - began).count();
    const double seconds = stepping;  // as the reference's GPU path reports it (main.rs:153-157)

    std::string fields_json = "{";
    for (size_t k = 0; k < totals.size(); ++k)
        fields_json += (k ? ", " : "") + std::string("\"") + bp.fields[k].name + "\": " + fmt_num(totals[k]);
    fields_json += "}";
    std::string summary = "{\n";
    summary += "  \"name\": \"" + bp.name + "\",\n";
    summary += "  \"engine\": \"" + engine + "\",\n";
    summary += "  \"steps\": " + std::to_string(bp.steps) + ",\n";
    summary += "  \"seconds\": " + fmt_num(seconds) + ",\n";
    summary += "  \"setup_seconds\": " + fmt_num(setup) + ",\n";
    summary += "  \"overall_seconds\": " + fmt_num(overall) + ",\n";
    summary += "  \"n_cells\": " + std::to_string(live) + ",\n";
    summary += "  \"n_labels\": " + std::to_string(vol.size() ? vol.size() - 1 : 0) + ",\n";
    summary += "  \"fields\": " + fields_json + ",\n";
    summary += "  \"mean_volume\": " + fmt_num(mean) + ",\n";
    summary += "  \"width\": " + std::to_string(bp.width) + ",\n";
    summary += "  \"height\": " + std::to_string(bp.height) + ",\n";
    summary += "  \"micron_per_site\": " + fmt_num(bp.micron_per_site) + ",\n";
    summary += "  \"minute_per_step\": " + fmt_num(bp.minute_per_step) + ",\n";
    const double step_time = stepping - setup > 0.0 ? stepping - setup : stepping;
    summary += "  \"site_attempts_per_second\": " + fmt_num((double)attempted / step_time) + ",\n";
    summary += "  \"acceptance\": " + fmt_num(attempted ? (double)accepted / (double)attempted : 0.0) + "\n";
    summary += "}";
    {
        std::ofstream s(join(out_dir, "summary.json"));
        if (!s) {
            std::fprintf(stderr, "cannot write the summary\n");
            return 1;
        }
        s << summary;
    }
    if (!npy_write_u32(join(out_dir, "labels_" + pad5(bp.steps) + ".npy"), lab.data(), m.width, m.height,
                       m.depth, &err)) {
        std::fprintf(stderr, "cannot write the final field: %s\n", err.c_str());
        return 1;
    }
    std::printf("%s: %llu steps in %.3f s, %zu cells, mean volume %.1f\n", engine.c_str(),
                (unsigned long long)bp.steps, seconds, live, mean);
    ghip_destroy(sim);
    return 0;
}
